"""Ray-parallel trace (march grid with out-of-bounds border bits, LDS ray ring,
per-ray result slots, per-pixel resolve) bit for bit against the oracle:
odd and even volume sizes smaller than the model (rays leave through every
face), every launch tuning, and after each kind of volume edit (full upload,
region upload, rebuild with a moved instance, clear)."""
import numpy as np
import pytest
import torch

from helpers import IDENTITY, assert_bitwise

pytestmark = pytest.mark.gpu

TUNINGS = [dict(kernel=2),
           dict(kernel=2, block=512, refill=1, steps=2),
           dict(kernel=2, refill=64, steps=8, tile=1),
           dict(kernel=2, refill=7, steps=4, blocks_per_cu=1, tile=2, order=1),
           dict(kernel=2, tile=0, order=1),
           dict(kernel=2, chunks=1),
           dict(kernel=2, chunks=5, tile=1),
           dict(kernel=2, chunks=64, order=1, refill=3),
           dict(kernel=2, block=512, steps=16, chunks=2)]


@pytest.fixture(scope="module")
def api():
    from voxelengine_amd import api as A
    return A


def _model(api, ctx, dims, seed):
    m = api.Model(ctx, *dims)
    m.generate(seed=seed, terrain_fill=0.25, nboxes=30, box_min=2, box_max=12)
    m.build_mips()
    return m


def _trace_both(api, ctx, oracle, occ, occ_h, model, variant, params, w=72, h=52):
    from voxelengine_amd import configs, views
    vd = views.scene_view(model.dims, w, h, frame=variant, variant=variant)
    g = api.GBuffer(w, h)
    api.primary(ctx, model, vd, g)
    pal = configs.make_palette()
    vis = api.Visibility(w, h, fetches=True)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
    ref = oracle.trace(occ_h, vd, params, g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(),
                       pal)
    got = (vis.shadow.cpu().numpy(), vis.ao.cpu().numpy(), vis.spec.cpu().numpy(),
           vis.fetches.cpu().numpy().astype(np.uint32))
    # the launch without fetch counts keeps a per-block step count: same bits
    vis2 = api.Visibility(w, h)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis2)
    for name, a, b in zip(("shadow", "ao", "spec"), (vis2.shadow, vis2.ao, vis2.spec), ref):
        assert_bitwise(a.cpu().numpy(), b, f"{name} without fetch counts")
    return got, ref


def _assert_same(got, ref, what):
    for name, a, b in zip(("shadow", "ao", "spec", "fetches"), got, ref):
        assert_bitwise(a, b, f"{what} {name}")


@pytest.mark.parametrize("vol_dims", [(45, 37, 43), (48, 40, 44), (17, 50, 9)])
def test_march_volume_sizes_and_tunings(api, ctx, oracle, vol_dims):
    from voxelengine_amd import configs
    model = _model(api, ctx, (96, 80, 88), seed=11)
    vol = model.download(0)
    occ = api.Volume(ctx, *vol_dims, nlevels=1)
    occ.clear()
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, IDENTITY)])
    occ.build_levels()
    occ_h = np.zeros(vol_dims[::-1], np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    assert_bitwise(occ.download(), occ_h, "occupancy")
    params = configs.trace_params(n_ao=5, ao_maxt=30.0, march="dda")
    try:
        for ti, tu in enumerate(TUNINGS):
            ctx.set_tuning(**tu)
            for variant in range(4):
                got, ref = _trace_both(api, ctx, oracle, occ, occ_h, model, variant, params)
                _assert_same(got, ref, f"dims {vol_dims} tuning {tu} view {variant}")
    finally:
        ctx.reset_tuning()


def test_step_limit_in_every_tuning(api, ctx, oracle):
    """Spec rays through a 320^3 volume holding a floor (y < 2) and a ceiling
    (y >= 150): a ray reflected off the floor needs more than 256 steps to
    reach the ceiling (300 fine voxels up), so many end at the 256-step limit
    (one ray per pixel, so such a pixel counts 4 * 256 fetches) and reads as a
    miss (spec 1), where a march that ran past the limit would report the
    ceiling's distance. Every tuning's block length, with and without fetch
    counts, matches the oracle."""
    from voxelengine_amd import configs
    model = _model(api, ctx, (96, 80, 88), seed=11)
    S = (320, 320, 320)
    host = np.zeros(S[::-1], np.uint8)
    host[:, :2, :] = 9                       # a floor: y < 2
    host[:, 150:, :] = 9                     # a ceiling: y >= 150
    occ = api.Volume(ctx, *S, nlevels=1)
    occ.upload_region(host, [(0, 0, 0, S[0], S[1], S[2])])
    params = configs.trace_params(n_shadow=0, n_ao=0, n_spec=1, spec_maxt=2000.0, march="dda")
    try:
        for tu in TUNINGS:
            ctx.set_tuning(**tu)
            for variant in (0, 2):
                got, ref = _trace_both(api, ctx, oracle, occ, host, model, variant, params)
                _assert_same(got, ref, f"tuning {tu} view {variant}")
                limit = ref[3] == 1024
                assert limit.sum() > 50
                assert (ref[2][limit] == 1.0).all()
    finally:
        ctx.reset_tuning()


def test_march_after_volume_edits(api, ctx, oracle):
    from voxelengine_amd import configs
    model = _model(api, ctx, (64, 64, 64), seed=12)
    vol = model.download(0)
    S = (32, 32, 32)
    params = configs.trace_params(n_ao=4, ao_maxt=24.0, march="dda")
    occ = api.Volume(ctx, *S, nlevels=2)
    rng = np.random.default_rng(3)

    # full upload of random texels
    host = (rng.random(S[::-1]) < 0.04).astype(np.uint8) * rng.integers(1, 256, S[::-1]).astype(np.uint8)
    occ.upload_region(host, [(0, 0, 0, S[0], S[1], S[2])])
    got, ref = _trace_both(api, ctx, oracle, occ, host, model, 0, params)
    _assert_same(got, ref, "full upload")

    # region upload over part of it
    host2 = host.copy()
    box = (5, 3, 7, 20, 11, 9)
    host2[box[2]:box[2] + box[5], box[1]:box[1] + box[4], box[0]:box[0] + box[3]] = 0xFF
    occ.upload_region(host2, [box])
    got, ref = _trace_both(api, ctx, oracle, occ, host2, model, 1, params)
    _assert_same(got, ref, "region upload")

    # rebuild from the model, then move the instance (erase old, write new)
    occ.clear()
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, IDENTITY)])
    occ_h = np.zeros(S[::-1], np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    moved = np.eye(4, dtype=np.float32)
    moved[3, :3] = (0.9, -0.6, 1.3)          # column-major translation (world units of 10 voxels)
    moved = moved.reshape(16)
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, moved)])
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, moved)])
    assert_bitwise(occ.download(), occ_h, "moved occupancy")
    got, ref = _trace_both(api, ctx, oracle, occ, occ_h, model, 2, params)
    _assert_same(got, ref, "moved instance")

    # clear: nothing occludes
    occ.clear()
    got, ref = _trace_both(api, ctx, oracle, occ, np.zeros(S[::-1], np.uint8), model, 3, params)
    _assert_same(got, ref, "cleared")
    hit = ref[0] > 0
    assert hit.any() and (ref[1][hit] == 1.0).all()


def test_march_profile_counts_launches(api, ctx):
    from voxelengine_amd import configs, views
    model = _model(api, ctx, (32, 32, 32), seed=13)
    occ = api.Volume(ctx, 16, 16, 16, nlevels=0)
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, IDENTITY)])
    vd = views.scene_view(model.dims, 40, 30)
    g = api.GBuffer(40, 30)
    api.primary(ctx, model, vd, g)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_ao=2, ao_maxt=8.0, march="dda")
    vis = api.Visibility(40, 30)
    ctx.profile(True)
    for _ in range(3):
        api.trace(ctx, occ, vd, g, pal, params, vis)
    ms, n = ctx.profile_read()
    ctx.profile(False)
    assert n == 3 and ms > 0.0


def test_no_rays_one_ray_class_and_tiny_views(api, oracle):
    """Zero rays per pixel (no march: k_resolve writes lit / 1 / 1), one ray
    class at a time, and 1x1, sub-tile and ragged views, starting on a fresh
    context whose trace buffers do not exist yet; then the same through a
    batch."""
    from voxelengine_amd import configs
    ctx = api.Context(0)
    model = _model(api, ctx, (40, 32, 36), seed=5)
    vol = model.download(0)
    dims = (20, 16, 18)
    occ = api.Volume(ctx, *dims, nlevels=1)
    occ.clear()
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, IDENTITY)])
    occ.build_levels()
    occ_h = np.zeros(dims[::-1], np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    for counts in [(0, 0, 0), (0, 3, 0), (1, 0, 0), (0, 0, 1)]:
        params = configs.trace_params(n_shadow=counts[0], n_ao=counts[1], n_spec=counts[2], ao_maxt=20.0,
                                      march="dda")
        for w, h in [(1, 1), (7, 3), (16, 9)]:
            got, ref = _trace_both(api, ctx, oracle, occ, occ_h, model, 1, params, w, h)
            _assert_same(got, ref, f"rays {counts} view {w}x{h}")
            if counts == (0, 0, 0):
                # hit pixels read lit / 1 / 1, sky pixels 0 / 0 / 0
                lit = got[0] == 1
                assert np.array_equal(got[1] == 1.0, lit) and np.array_equal(got[2] == 1.0, lit)
                assert not got[3].any()
    # a batch without rays, and a batch of tiny views
    from voxelengine_amd import views
    pal = torch.from_numpy(configs.make_palette()).cuda()
    for counts in [(0, 0, 0), (1, 2, 1)]:
        params = configs.trace_params(n_shadow=counts[0], n_ao=counts[1], n_spec=counts[2], ao_maxt=20.0,
                                      march="dda")
        items, plain = [], []
        for k, (w, h) in enumerate([(1, 1), (7, 3), (16, 9), (3, 17)]):
            vd = views.scene_view(model.dims, w, h, frame=k, variant=k)
            g = api.GBuffer(w, h)
            api.primary(ctx, model, vd, g)
            p = api.Visibility(w, h, fetches=True)
            api.trace(ctx, occ, vd, g, pal, params, p)
            items.append((occ, vd, g, api.Visibility(w, h, fetches=True)))
            plain.append(p)
        batch = api.Batch(ctx, items, pal, params)
        batch.launch()
        torch.cuda.synchronize()
        for k, ((_, _, _, o), p) in enumerate(zip(items, plain)):
            for f in ("shadow", "ao", "spec", "fetches"):
                assert_bitwise(getattr(o, f).cpu().numpy(), getattr(p, f).cpu().numpy(), f"batch {counts} item {k} {f}")
        batch.close()


@pytest.mark.parametrize("direction", [(1, 1, 1), (-1, 1, -1), (1, 1, 2), (2, -1, -1)])
def test_ties_match_oracle(api, ctx, oracle, direction):
    """Shadow rays from voxel centres along directions with equal components,
    so that two or three t_max tie at every step (the diagonal steps, their
    doubled and tripled best_t against maxt, and the step count), bit for bit
    against the oracle, at the default and at a short maxt."""
    from voxelengine_amd import configs, views
    rng = np.random.default_rng(41)
    dims = (40, 40, 40)
    occ_h = ((rng.random(dims[::-1]) < 0.04).astype(np.uint8) *
             rng.integers(1, 256, dims[::-1]).astype(np.uint8))
    occ = api.Volume(ctx, *dims, nlevels=1)
    occ.clear()
    occ.upload_region(occ_h, [(0, 0, 0, *dims)])
    occ.build_levels()
    w, h = 32, 32
    vd = views.scene_view((80, 80, 80), w, h)
    g = api.GBuffer(w, h)
    cells = rng.integers(2, 76, (w * h, 3)).astype(np.float32) + 0.5      # fine-voxel centres
    g.pos.copy_(torch.from_numpy(cells))
    g.normal.zero_()                                                     # the origin is the centre itself
    g.material.fill_(20)
    d = np.asarray(direction, np.float32)
    d = (d / np.float32(np.sqrt(np.float32((d * d).sum())))).astype(np.float32)
    pal = configs.make_palette()
    for maxt in (128.0, 7.3):
        params = configs.trace_params(n_shadow=1, n_ao=0, n_spec=0, shadow_maxt=maxt, march="dda", jitter_shadow=0)
        params["light_dir"] = [float(x) for x in d]
        vis = api.Visibility(w, h, fetches=True)
        api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
        sh, _, _, fe = oracle.trace(occ_h, vd, params, cells, np.zeros((w * h, 3), np.float32),
                                    np.full(w * h, 20, np.uint8), pal)
        assert_bitwise(vis.shadow.cpu().numpy(), sh, f"shadow {direction} maxt {maxt}")
        assert_bitwise(vis.fetches.cpu().numpy().astype(np.uint32), fe, f"fetches {direction} maxt {maxt}")
        assert 0 < sh.sum() < sh.size
