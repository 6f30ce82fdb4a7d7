"""The workloads P, B and L on the device, against plain launches and the
oracle.

P (256^3 terrain, 640x360, 1 shadow + 4 AO rays capped at 16, the reference's
marchers): every pixel against the oracle, from the scene generator to the
visibility. B (64 scenes of uneven size, one view each, 1 shadow + 8 AO): the
batched graph launch equals the 64 plain launches bit for bit on every item,
and four whole items equal the oracle. L (1024^3 city, four 3840x2160 views,
1 shadow + 32 AO + 1 spec): the batched launch equals the plain launches on
all four views, and one view equals the oracle on every 16th 8x8 tile."""
import numpy as np
import pytest
import torch

from helpers import IDENTITY, assert_bitwise

pytestmark = pytest.mark.gpu


def _device_scene(api, ctx, dims, seed, fill, nboxes, box, grounded=False, levels=2):
    m = api.Model(ctx, *dims)
    m.generate(seed=seed, terrain_fill=fill, nboxes=nboxes, box_min=box[0], box_max=box[1], grounded=grounded)
    m.build_mips()
    occ = api.Volume(ctx, dims[0] // 2, dims[1] // 2, dims[2] // 2, nlevels=levels)
    occ.clear()
    occ.rebuild([api.make_instance(m.mip_ptr(0), m.dims, IDENTITY, IDENTITY)])
    occ.build_levels()
    return m, occ


def _vis_np(vis):
    return (vis.shadow.cpu().numpy(), vis.ao.cpu().numpy(), vis.spec.cpu().numpy(),
            vis.fetches.cpu().numpy().astype(np.uint32))


def _assert_vis(got, ref, what, pixels=None):
    for name, a, b in zip(("shadow", "ao", "spec", "fetches"), got, ref):
        if pixels is not None:
            a, b = a[pixels], b[pixels]
        assert_bitwise(a, b, f"{what} {name}")


@pytest.mark.parametrize("march", ["reference", "dda"])
def test_config_P_every_pixel(ctx, oracle, march):
    from voxelengine_amd import api, configs
    cfg = configs.CONFIGS["P"]
    dims, W, H = cfg["dims"], cfg["width"], cfg["height"]
    m, occ = _device_scene(api, ctx, dims, configs.SEED, cfg["fill"], cfg["nboxes"], cfg["box"], cfg["grounded"])
    vol = m.download(0)
    ref_vol = oracle.gen_terrain(*dims, configs.SEED, cfg["fill"])
    assert_bitwise(vol, ref_vol, "P scene")
    assert 0.25 < (vol > 0).mean() < 0.35
    occ_h = np.zeros((dims[2] // 2, dims[1] // 2, dims[0] // 2), np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    assert_bitwise(occ.download(), occ_h, "P occupancy")
    vd = configs.config_view(cfg)
    g = api.GBuffer(W, H)
    api.primary(ctx, m, vd, g)
    pos, nrm, mat = oracle.primary(oracle.build_mips(vol), vd)
    assert_bitwise(g.material.cpu().numpy(), mat, "P material")
    assert_bitwise(g.pos.cpu().numpy(), pos, "P pos")
    assert_bitwise(g.normal.cpu().numpy(), nrm, "P normal")
    params = configs.trace_params(**dict(cfg["params"], march=march))
    pal = configs.make_palette()
    vis = api.Visibility(W, H, fetches=True)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
    ref = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal)
    _assert_vis(_vis_np(vis), ref, f"P {march}")
    hit = mat > 0
    assert hit.mean() > 0.5 and (ref[0][hit] == 0).any() and (ref[1][hit] < 1).any()


def test_config_B_batched_equals_plain_and_oracle(ctx, oracle):
    from voxelengine_amd import api, configs, views
    params = configs.trace_params(n_shadow=1, n_ao=8, ao_maxt=32.0, n_spec=0, march="dda")
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    scenes = configs.batched_scenes()
    assert len(scenes) == 64
    items, plain, models = [], [], []
    for sc in scenes:
        m, occ = _device_scene(api, ctx, sc["dims"], sc["seed"], sc["fill"], sc["nboxes"], sc["box"])
        vd = views.scene_view(sc["dims"], sc["width"], sc["height"], variant=sc["variant"])
        g = api.GBuffer(sc["width"], sc["height"])
        api.primary(ctx, m, vd, g)
        p = api.Visibility(sc["width"], sc["height"], fetches=True)
        api.trace(ctx, occ, vd, g, pal_d, params, p)
        items.append((occ, vd, g, api.Visibility(sc["width"], sc["height"], fetches=True)))
        plain.append(p)
        models.append(m)
    batch = api.Batch(ctx, items, pal_d, params)
    batch.launch()
    torch.cuda.synchronize()
    for k, ((_, _, _, b), p) in enumerate(zip(items, plain)):
        _assert_vis(_vis_np(b), _vis_np(p), f"B item {k} batched vs plain")
    # replays of the graph give the same bits again
    batch.launch()
    torch.cuda.synchronize()
    for k in (0, 31, 63):
        _assert_vis(_vis_np(items[k][3]), _vis_np(plain[k]), f"B item {k} replay")
    batch.close()
    # four whole items against the oracle: the smallest and largest view, two others
    npx = [sc["width"] * sc["height"] for sc in scenes]
    picks = sorted({int(np.argmin(npx)), int(np.argmax(npx)), 17, 40})
    for k in picks:
        sc, (occ, vd, g, _), m = scenes[k], items[k], models[k]
        vol = m.download(0)
        d = sc["dims"]
        occ_h = np.zeros((d[2] // 2, d[1] // 2, d[0] // 2), np.uint8)
        oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
        assert_bitwise(occ.download(), occ_h, f"B item {k} occupancy")
        mips = [vol, m.download(1), m.download(2)]
        pos, nrm, mat = oracle.primary(mips, vd)
        assert_bitwise(g.material.cpu().numpy(), mat, f"B item {k} material")
        assert_bitwise(g.pos.cpu().numpy(), pos, f"B item {k} pos")
        ref = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal)
        _assert_vis(_vis_np(plain[k]), ref, f"B item {k} oracle")
    del items, plain, models
    torch.cuda.empty_cache()


def test_config_L_batched_equals_plain_and_oracle_subsample(ctx, oracle):
    from voxelengine_amd import api, configs
    cfg = configs.CONFIGS["L"]
    dims, W, H = cfg["dims"], cfg["width"], cfg["height"]
    params = configs.trace_params(**cfg["params"])
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    m, occ = _device_scene(api, ctx, dims, configs.SEED, cfg["fill"], cfg["nboxes"], cfg["box"], cfg["grounded"],
                           cfg["levels"])
    items, plain = [], []
    for k in range(cfg["nviews"]):
        vd = configs.config_view(cfg, variant=k)
        g = api.GBuffer(W, H)
        api.primary(ctx, m, vd, g)
        p = api.Visibility(W, H)
        api.trace(ctx, occ, vd, g, pal_d, params, p)
        items.append((occ, vd, g, api.Visibility(W, H)))
        plain.append(p)
    batch = api.Batch(ctx, items, pal_d, params)
    batch.launch()
    torch.cuda.synchronize()
    hit_frac = []
    for k, ((_, _, g, b), p) in enumerate(zip(items, plain)):
        for f in ("shadow", "ao", "spec"):
            assert torch.equal(getattr(b, f).view(torch.uint8), getattr(p, f).view(torch.uint8)), (k, f)
        hit_frac.append(float((g.material > 0).float().mean().item()))
    batch.close()
    assert min(hit_frac) > 0.9, hit_frac           # the views look into the city
    # view 0 against the oracle on every 16th 8x8 tile
    _, vd, g, _ = items[0]
    tiles_x = W // 8
    tiles = np.arange(0, tiles_x * (H // 8), 16)
    ty, tx = tiles // tiles_x, tiles % tiles_x
    yy = (ty[:, None] * 8 + np.arange(8)[None, :])
    xx = (tx[:, None] * 8 + np.arange(8)[None, :])
    pix = (yy[:, :, None] * W + xx[:, None, :]).reshape(-1).astype(np.int32)
    vol = m.download(0)
    mips = [vol, m.download(1), m.download(2)]
    pos_s, nrm_s, mat_s = oracle.primary(mips, vd, pixels=pix)
    pos, nrm, mat = g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy()
    assert_bitwise(mat[pix], mat_s, "L material")
    assert_bitwise(pos[pix], pos_s, "L pos")
    assert_bitwise(nrm[pix], nrm_s, "L normal")
    occ_h = np.zeros((dims[2] // 2, dims[1] // 2, dims[0] // 2), np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    assert_bitwise(occ.download(), occ_h, "L occupancy")
    assert 0.045 < (vol > 0).mean() < 0.06
    del vol, mips
    sh, ao, sp, fe = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal, pixels=pix)
    got = plain[0]
    assert_bitwise(got.shadow.cpu().numpy()[pix], sh[pix], "L shadow")
    assert_bitwise(got.ao.cpu().numpy()[pix], ao[pix], "L ao")
    assert_bitwise(got.spec.cpu().numpy()[pix], sp[pix], "L spec")
    assert (mat_s > 0).mean() > 0.9
    del items, plain, occ, m
    torch.cuda.empty_cache()
