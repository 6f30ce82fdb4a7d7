"""LightTAA on the device (k_light_taa) against the oracle, bit for bit, over
three frames of a moving camera with real geometry (the FarmHouse draws
through the geometry pass), noisy light and the history fed back; the last
frame jumps the camera so that many pixels take the off-screen branch."""
import numpy as np
import pytest

from helpers import assert_bitwise
from voxelengine_amd import views

pytestmark = pytest.mark.gpu


def test_taa_frames_match_oracle(ctx, oracle, tmp_path):
    import torch
    from golden_mods import write_mod_folder
    from voxelengine_amd import api, configs
    mods = write_mod_folder(str(tmp_path))
    P = api.Prefab(mods, "default/FarmHouse.pf")
    models = P.load_models(ctx)
    cmds = P.vox_cmds(models)
    W, H = 160, 96
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    eyes = [(8.0, 16.0, 6.0), (8.3, 16.1, 6.2), (8.6, 16.2, 6.4), (14.0, 12.0, 2.0)]
    rng = np.random.default_rng(5)
    last = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
    prev_view = None
    ood = 0
    for f, eye in enumerate(eyes):
        vd = views.make_view(eye, (32.0, 0.0, 32.0), W, H, frame=f)
        if prev_view is not None:
            vd["last_view"] = prev_view
        prev_view = vd["view"]
        g = api.GBuffer(W, H)
        ext = api.GBufferExt(W, H)
        api.primary_instances(ctx, cmds, vd, g, ext)
        light = torch.from_numpy((rng.random((W * H, 3)) * 2.0).astype(np.float32)).cuda()
        out = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
        api.light_taa(ctx, vd, light, last, out, ext.depth, g.normal, g.material, pal_d, motion=ext.motion, seed=9)
        torch.cuda.synchronize()
        ref = oracle.light_taa(W, H, f, 9, light.cpu().numpy(), last.cpu().numpy(), ext.motion.cpu().numpy(),
                               ext.depth.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(), pal)
        assert_bitwise(out.cpu().numpy(), ref, f"frame {f}")
        hit = g.material.cpu().numpy() > 0
        assert hit.mean() > 0.3
        mo = ext.motion.cpu().numpy()
        u = (np.arange(W * H) % W + 0.5) / W + mo[:, 0]
        v = (np.arange(W * H) // W + 0.5) / H + mo[:, 1]
        ood = max(ood, int((hit & ((u < 0) | (u > 1) | (v < 0) | (v > 1))).sum()))
        last = out
    assert ood > 50     # the jump sends pixels through the off-screen branch


def test_taa_extreme_motion_matches_oracle(ctx, oracle):
    """History positions far off the image, infinite or NaN (the motion of a
    draw whose last position projects behind the camera) read the border,
    the same on the device as in the oracle."""
    import torch
    from voxelengine_amd import api, configs
    W, H = 96, 64
    npx = W * H
    rng = np.random.default_rng(3)
    n = rng.normal(size=(npx, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    mat = rng.integers(1, 4, npx).astype(np.uint8)
    mat[rng.random(npx) < 0.1] = 0
    depth = np.where(mat > 0, rng.uniform(0.001, 0.02, npx), 1.0).astype(np.float32)
    motion = (rng.normal(size=(npx, 2)) * 0.01).astype(np.float32)
    special = np.array([np.inf, -np.inf, np.nan, 1e30, -1e30, 3e9, -3e9, 0.6, -0.6], np.float32)
    pick = rng.random(npx) < 0.3
    motion[pick, rng.integers(0, 2, pick.sum())] = special[rng.integers(0, len(special), pick.sum())]
    light = rng.random((npx, 3)).astype(np.float32)
    last = rng.random((npx, 4)).astype(np.float32)
    pal = configs.make_palette()
    vd = views.make_view((0.0, 2.0, 0.0), (1.0, 0.0, 1.0), W, H, frame=2)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = torch.zeros((npx, 4), dtype=torch.float32, device="cuda")
    api.light_taa(ctx, vd, d(light), d(last), out, d(depth), d(n), d(mat), d(pal), motion=d(motion), seed=4)
    torch.cuda.synchronize()
    ref = oracle.light_taa(W, H, 2, 4, light, last, motion, depth, n, mat, pal)
    assert_bitwise(out.cpu().numpy(), ref, "taa")
