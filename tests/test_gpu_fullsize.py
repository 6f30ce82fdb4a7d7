"""Full-size parity: the benchmark workload (512^3 terrain + 2000 boxes, 1920x1080,
1 shadow + 16 AO + 1 spec ray per pixel) traced by the product and by the oracle
(OpenMP on the host) over every pixel; plus size-independent properties."""
import numpy as np
import pytest
import torch

from helpers import IDENTITY, assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frame_f(ctx):
    from voxelengine_amd import api, configs, views
    cfg = configs.CONFIGS["F"]
    nx, ny, nz = cfg["dims"]
    model = api.Model(ctx, nx, ny, nz)
    model.generate(seed=configs.SEED, terrain_fill=cfg["fill"], nboxes=cfg["nboxes"], box_min=cfg["box"][0],
                   box_max=cfg["box"][1], grounded=cfg["grounded"])
    model.build_mips()
    occ = api.Volume(ctx, nx // 2, ny // 2, nz // 2, nlevels=cfg["levels"])
    occ.clear()
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, IDENTITY)])
    occ.build_levels()
    W, H = cfg["width"], cfg["height"]
    vd = views.scene_view(cfg["dims"], W, H)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, vd, g)
    torch.cuda.synchronize()
    return dict(cfg=cfg, model=model, occ=occ, vd=vd, g=g)


def test_fullsize_occupancy_and_primary(frame_f, oracle):
    from voxelengine_amd import configs
    vol = frame_f["model"].download(0)
    ref_vol = oracle.gen_terrain(512, 512, 512, configs.SEED, 0.30)
    oracle.gen_boxes(ref_vol, configs.SEED, 2000, 2, 32, 0)
    assert_bitwise(vol, ref_vol, "scene")
    occ_h = np.zeros((256, 256, 256), np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    assert_bitwise(frame_f["occ"].download(), occ_h, "occupancy")
    pos, nrm, mat = oracle.primary(oracle.build_mips(vol), frame_f["vd"])
    g = frame_f["g"]
    assert_bitwise(g.material.cpu().numpy(), mat, "material")
    assert_bitwise(g.pos.cpu().numpy(), pos, "pos")
    assert_bitwise(g.normal.cpu().numpy(), nrm, "normal")
    frame_f["occ_h"] = occ_h


@pytest.mark.parametrize("march", ["dda", "reference"])
def test_fullsize_trace_every_pixel(frame_f, oracle, ctx, march):
    from voxelengine_amd import api, configs
    cfg = frame_f["cfg"]
    params = configs.trace_params(**dict(cfg["params"], march=march))
    g, vd = frame_f["g"], frame_f["vd"]
    W, H = cfg["width"], cfg["height"]
    pal = configs.make_palette()
    vis = api.Visibility(W, H, fetches=True)
    api.trace(ctx, frame_f["occ"], vd, g, torch.from_numpy(pal).cuda(), params, vis)
    occ_h = frame_f.get("occ_h")
    if occ_h is None:
        occ_h = frame_f["occ"].download()
    pos, nrm, mat = g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy()
    sh, ao, sp, fe = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal)
    assert_bitwise(vis.shadow.cpu().numpy(), sh, "shadow")
    assert_bitwise(vis.ao.cpu().numpy(), ao, "ao")
    assert_bitwise(vis.spec.cpu().numpy(), sp, "spec")
    assert_bitwise(vis.fetches.cpu().numpy().astype(np.uint32), fe, "fetches")
    # repeat launch without the fetch counts: identical bits (AO/spec can hold NaN
    # where a zero-component ray hits its origin voxel, best_t = inf*0: compare bits)
    vis2 = api.Visibility(W, H, fetches=False)
    api.trace(ctx, frame_f["occ"], vd, g, torch.from_numpy(pal).cuda(), params, vis2)
    assert_bitwise(vis2.ao.cpu().numpy(), ao, "ao rerun")
    assert_bitwise(vis2.shadow.cpu().numpy(), sh, "shadow rerun")
    assert_bitwise(vis2.spec.cpu().numpy(), sp, "spec rerun")


def test_fullsize_hit_records_every_ray(frame_f, oracle, ctx):
    """Hit voxel, hit face (normal), best_t, step count and fetch count of all
    27.1M rays of the frame from the headline kernel (vx_trace_hits), against
    the oracle's raycastShadowVolume on the same rays (Light.frag:29-81)."""
    from voxelengine_amd import api, configs
    cfg = frame_f["cfg"]
    params = configs.trace_params(**cfg["params"])
    g, vd = frame_f["g"], frame_f["vd"]
    W, H = cfg["width"], cfg["height"]
    pal = configs.make_palette()
    got = api.trace_hits(ctx, frame_f["occ"], vd, g, torch.from_numpy(pal).cuda(), params, api.Visibility(W, H))
    occ_h = frame_f.get("occ_h")
    if occ_h is None:
        occ_h = frame_f["occ"].download()
    pos, nrm, mat = g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy()
    hit_px = np.nonzero(mat)[0].astype(np.int32)
    assert not got[mat == 0].view(np.uint8).any()
    got = got[hit_px]
    rays = oracle.pixel_rays(vd, params, pos, nrm, mat, pal, hit_px)
    ref = oracle.trace_rays(occ_h, rays.reshape(-1, 7), 0).reshape(hit_px.size, -1)
    del rays
    for f in ("term", "steps", "fetches", "voxel", "t", "normal"):
        assert_bitwise(got[f], ref[f], f"F frame {f}")
    terms, counts = np.unique(ref["term"], return_counts=True)
    assert set(terms.tolist()) == {0, 1, 2} and counts.sum() == 27093780
