"""Gradients of the shaded image (config D at reduced resolution and full size):
torch.autograd over the C ABI against FP64 central finite differences of the
oracle's FP64 forward and against the oracle's FP64 analytic adjoint.
Tolerance: 1e-4 relative."""
import numpy as np
import pytest
import torch

from helpers import IDENTITY, device_scene

pytestmark = pytest.mark.gpu


def _setup(ctx, N, W, H, seed=21):
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, N, seed=seed, fill=0.3, nboxes=0)
    vd = views.scene_view(model.dims, W, H, variant=3)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, vd, g)
    pal = configs.make_palette()
    params = configs.trace_params(n_ao=8, ao_maxt=32.0, n_spec=0)
    vis = api.Visibility(W, H)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
    return vd, g, vis, pal


def test_autograd_matches_fp64_finite_differences(ctx, oracle):
    from voxelengine_amd import autograd, configs
    W, H = 96, 64
    vd, g, vis, pal = _setup(ctx, 64, W, H)
    light = configs.default_light()
    light[3:] = [1.2, 1.0, 0.8]
    pal_t = torch.from_numpy(pal).cuda().requires_grad_(True)
    light_t = torch.from_numpy(light).requires_grad_(True)
    target_pal = configs.perturbed_palette(pal)
    host = [g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(), vis.shadow.cpu().numpy(),
            vis.ao.cpu().numpy()]
    target = oracle.shade_f64(vd, *host, target_pal.astype(np.float64), light.astype(np.float64))
    rgb = autograd.shade(pal_t, light_t, ctx, vd, g, vis)
    loss = ((rgb.double() - torch.from_numpy(target).cuda()) ** 2).sum()
    loss.backward()
    gp = pal_t.grad.cpu().numpy().astype(np.float64)
    gl = light_t.grad.numpy().astype(np.float64)

    def f64_loss(p, l):
        I = oracle.shade_f64(vd, *host, p, l)
        return ((I - target) ** 2).sum()

    pal64, light64 = pal.astype(np.float64), light.astype(np.float64)
    used = np.unique(host[2][host[2] > 0])
    rng = np.random.default_rng(0)
    h = 1e-6
    for m in rng.choice(used, size=min(10, used.size), replace=False):
        for k in range(6):
            p1, p2 = pal64.copy(), pal64.copy()
            p1[m, k] += h
            p2[m, k] -= h
            fd = (f64_loss(p1, light64) - f64_loss(p2, light64)) / (2 * h)
            assert abs(gp[m, k] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(gp).max()), (m, k, gp[m, k], fd)
    for k in range(6):
        l1, l2 = light64.copy(), light64.copy()
        l1[k] += h
        l2[k] -= h
        fd = (f64_loss(pal64, l1) - f64_loss(pal64, l2)) / (2 * h)
        assert abs(gl[k] - fd) <= 1e-4 * abs(fd), (k, gl[k], fd)


def test_fullsize_backward_matches_fp64_adjoint(ctx, oracle):
    """Config D: 256^3 terrain at 1280x720; every palette and light gradient."""
    from voxelengine_amd import api, configs
    cfg = configs.CONFIGS["D"]
    W, H = cfg["width"], cfg["height"]
    vd, g, vis, pal = _setup(ctx, 256, W, H, seed=configs.SEED)
    light = configs.default_light()
    host = [g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(), vis.shadow.cpu().numpy(),
            vis.ao.cpu().numpy()]
    target = oracle.shade_f64(vd, *host, configs.perturbed_palette(pal).astype(np.float64), light.astype(np.float64))
    rgb = torch.empty((W * H, 3), dtype=torch.float32, device="cuda")
    pal_d = torch.from_numpy(pal).cuda()
    api.shade_forward(ctx, vd, g, vis, pal_d, light, rgb)
    dl = (2.0 * (rgb.double() - torch.from_numpy(target).cuda())).float().contiguous()
    dpal = torch.empty((256, 6), dtype=torch.float32, device="cuda")
    dlight = torch.empty(6, dtype=torch.float32, device="cuda")
    api.shade_backward(ctx, vd, g, vis, pal_d, light, dl, dpal, dlight)
    rp, rl = oracle.shade_backward_f64(vd, *host, pal.astype(np.float64), light.astype(np.float64),
                                       dl.cpu().numpy().astype(np.float64))
    gp, gl = dpal.cpu().numpy().astype(np.float64), dlight.cpu().numpy().astype(np.float64)
    scale = np.abs(rp).max()
    assert np.abs(gp - rp).max() <= 1e-4 * scale
    big = np.abs(rp) > 1e-2 * scale
    assert (np.abs(gp - rp)[big] <= 1e-4 * np.abs(rp)[big]).all()
    assert np.abs(gl - rl).max() <= 1e-4 * np.abs(rl).max()


def test_batched_backward_accumulates_views(ctx, oracle):
    """Several views in one adjoint launch: one view alone equals vx_shade_backward
    bit for bit, a view given twice doubles it exactly (fixed-point sums), and two
    different views match the sum of the FP64 adjoints within 1e-4."""
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, 64, seed=23, fill=0.3, nboxes=10)
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    light = configs.default_light()
    params = configs.trace_params(n_ao=4, ao_maxt=16.0, n_spec=0)
    rng = np.random.default_rng(9)
    items, host = [], []
    for k, (w, h) in enumerate([(80, 48), (57, 71)]):
        vd = views.scene_view(model.dims, w, h, variant=k)
        g = api.GBuffer(w, h)
        api.primary(ctx, model, vd, g)
        vis = api.Visibility(w, h)
        api.trace(ctx, occ, vd, g, pal_d, params, vis)
        dl = torch.from_numpy(rng.normal(size=(w * h, 3)).astype(np.float32)).cuda()
        items.append((vd, g, vis, dl, None))
        host.append((vd, [g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(),
                          vis.shadow.cpu().numpy(), vis.ao.cpu().numpy()], dl.cpu().numpy().astype(np.float64)))

    def grads(fn):
        dp = torch.zeros((256, 6), dtype=torch.float32, device="cuda")
        dlt = torch.zeros(6, dtype=torch.float32, device="cuda")
        fn(dp, dlt)
        return dp.cpu().numpy(), dlt.cpu().numpy()

    vd0, g0, vis0, dl0, _ = items[0]
    sp, sl = grads(lambda dp, dlt: api.shade_backward(ctx, vd0, g0, vis0, pal_d, light, dl0, dp, dlt))
    bp, bl = grads(lambda dp, dlt: api.shade_backward_batch(ctx, items[:1], pal_d, light, dp, dlt))
    assert np.array_equal(sp.view(np.uint32), bp.view(np.uint32)) and np.array_equal(sl.view(np.uint32),
                                                                                    bl.view(np.uint32))
    dp2, dl2 = grads(lambda dp, dlt: api.shade_backward_batch(ctx, [items[0], items[0]], pal_d, light, dp, dlt))
    assert np.array_equal(dp2, 2.0 * sp) and np.array_equal(dl2, 2.0 * sl)
    gp, gl = grads(lambda dp, dlt: api.shade_backward_batch(ctx, items, pal_d, light, dp, dlt))
    rp = np.zeros((256, 6))
    rl = np.zeros(6)
    for vd, hst, dl in host:
        p, l = oracle.shade_backward_f64(vd, *hst, pal.astype(np.float64), light.astype(np.float64), dl)
        rp += p
        rl += l
    assert np.abs(gp - rp).max() <= 1e-4 * np.abs(rp).max()
    assert np.abs(gl - rl).max() <= 1e-4 * np.abs(rl).max()
