"""Point/spot light pass on the GPU against the oracle, bit for bit: rgb
accumulated over lights in order and the marchers' fetch counts, with the
reference marchers (Sparse for point, SuperSparse for spot) and with the DDA,
through both kernels of the pass (the (pixel, light) ray list, one ray per
lane; and the per-pixel light loop)."""
import numpy as np
import pytest
import torch

from helpers import assert_bitwise, device_scene, oracle_occ

pytestmark = pytest.mark.gpu


def _lights(dims, n_point=8, n_spot=5, seed=4):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    L = np.zeros(n_point + n_spot, O.LIGHT_DTYPE)
    ext = np.array(dims, np.float32) * 0.1          # world extent of the model
    for k in range(L.size):
        L[k]["type"] = 0 if k < n_point else 1
        L[k]["position"] = rng.uniform(-0.2, 1.2, 3) * ext
        L[k]["range"] = rng.uniform(0.3, 1.2) * ext.max()
        L[k]["color"] = rng.uniform(0.2, 4.0, 3)
        L[k]["attenuation"] = rng.uniform(0.5, 2.0)
        d = rng.normal(size=3)
        L[k]["direction"] = d / np.linalg.norm(d)
        L[k]["angle"] = rng.uniform(0.2, 0.9)
        L[k]["angle_attenuation"] = rng.uniform(0.5, 3.0)
    return L


@pytest.fixture(params=[1, 0], ids=["ray_list", "pixel_loop"])
def lights_path(request, ctx):
    ctx.set_tuning(lights=request.param)
    yield request.param
    ctx.reset_tuning()


@pytest.mark.parametrize("marchers", [(1, 2), (0, 0)])
def test_light_pass_matches_oracle(ctx, oracle, marchers, lights_path):
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, 64, seed=31, fill=0.3, nboxes=20)
    occ_h = oracle_occ(oracle, vol)
    w, h = 72, 50
    vd = views.scene_view(model.dims, w, h, variant=2)
    g = api.GBuffer(w, h)
    api.primary(ctx, model, vd, g)
    pal = configs.make_palette()
    L = _lights(model.dims)
    base = np.random.default_rng(5).random((w * h, 3)).astype(np.float32)
    rgb = torch.from_numpy(base.copy()).cuda()
    fetches = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    api.light_pass(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), api.lights_to_device(L), L.size, rgb,
                   point_march=marchers[0], spot_march=marchers[1], seed=17, frame=3, fetches=fetches)
    ref, ref_f = oracle.light_pass(occ_h, vd, L, oracle.light_params(marchers[0], marchers[1], 17, 3),
                                   g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(), pal,
                                   rgb=base)
    assert_bitwise(rgb.cpu().numpy(), ref, "rgb")
    assert_bitwise(fetches.cpu().numpy().astype(np.uint32), ref_f, "fetches")
    lit = (ref != base).any(axis=1)
    assert lit.mean() > 0.2 and (~lit & (g.material.cpu().numpy() > 0)).any()


def test_many_small_lights_match_oracle(ctx, oracle, lights_path):
    """150 lights with small ranges, so each 16x16 tile keeps only a few of them
    (the per-tile cull), plus ranges that are infinite, negative and zero:
    rgb and fetch counts bit for bit against the oracle, which tests every
    light at every pixel. (A NaN range, kept by the cull, makes the pixel's
    sum NaN, whose bits the two sides need not share; it is left out.)"""
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, 64, seed=32, fill=0.3, nboxes=20)
    occ_h = oracle_occ(oracle, vol)
    w, h = 120, 88
    vd = views.scene_view(model.dims, w, h, variant=1)
    g = api.GBuffer(w, h)
    api.primary(ctx, model, vd, g)
    pal = configs.make_palette()
    L = _lights(model.dims, n_point=80, n_spot=70, seed=9)
    rng = np.random.default_rng(10)
    L["range"] = rng.uniform(0.05, 0.25, L.size) * np.float32(model.dims[0] * 0.1)
    L["range"][[40, 90, 120]] = [np.inf, -1.0, 0.0]
    base = np.random.default_rng(11).random((w * h, 3)).astype(np.float32)
    rgb = torch.from_numpy(base.copy()).cuda()
    fetches = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    api.light_pass(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), api.lights_to_device(L), L.size, rgb,
                   seed=5, frame=2, fetches=fetches)
    ref, ref_f = oracle.light_pass(occ_h, vd, L, oracle.light_params(1, 2, 5, 2), g.pos.cpu().numpy(),
                                   g.normal.cpu().numpy(), g.material.cpu().numpy(), pal, rgb=base)
    got = rgb.cpu().numpy()
    assert_bitwise(got, ref, "rgb")
    assert_bitwise(fetches.cpu().numpy().astype(np.uint32), ref_f, "fetches")
    lit = (ref != base).any(axis=1)
    assert lit.mean() > 0.02


def test_no_light_in_range_and_no_lights(ctx, oracle, lights_path):
    """Lights whose ranges reach no hit pixel leave the light buffer as it was
    (the ray list is empty), and so does a pass with no lights."""
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, 32, seed=33, fill=0.3, nboxes=5)
    w, h = 40, 30
    vd = views.scene_view(model.dims, w, h)
    g = api.GBuffer(w, h)
    api.primary(ctx, model, vd, g)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    L = _lights(model.dims, n_point=3, n_spot=2)
    L["position"] = [1000.0, 1000.0, 1000.0]
    L["range"] = 1.0
    base = np.random.default_rng(2).random((w * h, 3)).astype(np.float32)
    for n in (L.size, 0):
        rgb = torch.from_numpy(base.copy()).cuda()
        fetches = torch.full((w * h,), 7, dtype=torch.int32, device="cuda")
        api.light_pass(ctx, occ, vd, g, pal, api.lights_to_device(L), n, rgb, fetches=fetches)
        torch.cuda.synchronize()
        assert_bitwise(rgb.cpu().numpy(), base, f"rgb with {n} lights")
        hit = g.material.cpu().numpy() > 0
        assert (fetches.cpu().numpy()[hit] == 0).all()
