"""Every per-pixel pass on views smaller than, and not a multiple of, its
16x16 tile (1x1, 7x3, 3x17, 17x1): the primary hit, its D24 depth, the
screen-space occlusion, the light pass on both paths and the reflection pass,
bit for bit against the oracle. The passes map pixels to tiles per view, so a
view that leaves most of a tile's lanes without a pixel must still write
exactly its own pixels."""
import numpy as np
import pytest
import torch

from helpers import assert_bitwise, device_scene, oracle_occ

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (7, 3), (3, 17), (17, 1)]


@pytest.fixture(scope="module")
def scene(ctx, oracle):
    from voxelengine_amd import api
    model, occ, vol = device_scene(api, ctx, 64, seed=23, fill=0.3, nboxes=25)
    return dict(model=model, occ=occ, occ_h=oracle_occ(oracle, vol), mips=oracle.build_mips(vol))


def _lights(dims):
    from voxelengine_amd import api
    rng = np.random.default_rng(9)
    ext = np.array(dims, np.float64) * 0.1
    L = np.zeros(12, api.LIGHT_DTYPE)
    for k in range(L.size):
        L[k]["type"] = k % 2
        L[k]["position"] = rng.uniform([0.0, 0.3, 0.0], [1.0, 0.8, 1.0]) * ext
        L[k]["range"] = rng.uniform(0.3, 0.8) * ext.max()
        L[k]["color"] = rng.uniform(0.5, 3.0, 3)
        L[k]["attenuation"] = rng.uniform(0.5, 2.0)
        d = rng.normal(size=3)
        L[k]["direction"] = d / np.linalg.norm(d)
        L[k]["angle"] = rng.uniform(0.3, 0.9)
        L[k]["angle_attenuation"] = rng.uniform(0.5, 2.0)
    return L


@pytest.mark.parametrize("size", SIZES, ids=[f"{w}x{h}" for w, h in SIZES])
def test_small_views_match_oracle(ctx, oracle, scene, size):
    from voxelengine_amd import api, configs, views
    W, H = size
    npx = W * H
    # a view centred on the scene, so that even a 1x1 view hits it
    vd = views.scene_view(scene["model"].dims, W, H, frame=1)
    g = api.GBuffer(W, H)
    api.primary(ctx, scene["model"], vd, g)
    depth = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.gbuf_depth(ctx, vd, g, depth, volume_rid=1)
    torch.cuda.synchronize()
    pos, nrm, mat = oracle.primary(scene["mips"], vd)
    assert_bitwise(g.pos.cpu().numpy(), pos, "pos")
    assert_bitwise(g.normal.cpu().numpy(), nrm, "normal")
    assert_bitwise(g.material.cpu().numpy(), mat, "material")
    d = depth.cpu().numpy()
    assert ((d >= 0) & (d <= 1)).all()

    # screen-space occlusion
    s = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.ssao(ctx, vd, g, depth, s, seed=5)
    torch.cuda.synchronize()
    assert_bitwise(s.cpu().numpy(), oracle.ssao(vd, nrm, mat, d, seed=5), "ssao")

    # light pass, both paths
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    L = _lights(scene["model"].dims)
    base = np.random.default_rng(3).random((npx, 3)).astype(np.float32)
    ref, ref_f = oracle.light_pass(scene["occ_h"], vd, L, oracle.light_params(1, 2, 11, 1), pos, nrm, mat, pal,
                                   rgb=base)
    for path in (0, 1):
        ctx.set_tuning(lights=path)
        try:
            rgb = torch.from_numpy(base.copy()).cuda()
            fetches = torch.full((npx,), -1, dtype=torch.int32, device="cuda")
            api.light_pass(ctx, scene["occ"], vd, g, pal_d, api.lights_to_device(L), L.size, rgb, point_march=1,
                           spot_march=2, seed=11, frame=1, fetches=fetches)
            torch.cuda.synchronize()
        finally:
            ctx.reset_tuning()
        assert_bitwise(rgb.cpu().numpy(), ref, f"light rgb, path {path}")
        assert_bitwise(fetches.cpu().numpy().astype(np.uint32), ref_f, f"light fetches, path {path}")

    # reflection pass
    light = np.random.default_rng(4).uniform(0, 3, (npx, 4)).astype(np.float32)
    out = torch.full((npx, 4), -1.0, dtype=torch.float32, device="cuda")
    t = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.reflection(ctx, scene["occ"], vd, g, pal_d, torch.from_numpy(light).cuda(), depth, out, t=t, march=1,
                   maxt=256.0, seed=2, frame=1)
    torch.cuda.synchronize()
    rref, tref = oracle.reflection(scene["occ_h"], vd, oracle.reflection_params(1, 256.0, 2, 1), pos, nrm, mat, pal,
                                   d, light)
    assert_bitwise(t.cpu().numpy(), tref, "reflection t")
    assert_bitwise(out.cpu().numpy(), rref, "reflection out")


def test_one_pixel_view_hits_the_scene(ctx, scene):
    """The 1x1 view above is not vacuous: its pixel hits geometry."""
    from voxelengine_amd import api, views
    vd = views.scene_view(scene["model"].dims, 1, 1, frame=1)
    g = api.GBuffer(1, 1)
    api.primary(ctx, scene["model"], vd, g)
    torch.cuda.synchronize()
    assert int(g.material.cpu().numpy()[0]) != 0
