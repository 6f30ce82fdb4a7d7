"""Hit voxel, hit face and step count of every ray the ray-parallel trace
kernel marches (through vx_trace_hits) against the oracle's
raycastShadowVolume (Light.frag:29-81) on the same rays: scenes that rays
leave through every face, every launch tuning, exact ties, and rays that end
at the 256-step limit below an occupied ceiling they never reach."""
import numpy as np
import pytest
import torch

from helpers import IDENTITY, assert_bitwise

pytestmark = pytest.mark.gpu

TUNINGS = [dict(kernel=2),
           dict(kernel=2, block=512, refill=1, steps=2, tile=2),
           dict(kernel=2, refill=64, steps=8, tile=1, order=1),
           dict(kernel=2, refill=7, steps=4, blocks_per_cu=1, order=1),
           dict(kernel=2, chunks=1),
           dict(kernel=2, chunks=4, tile=2),
           dict(kernel=2, block=512, steps=16, refill=32)]


def oracle_hits(oracle, occ_h, vd, params, pos, nrm, mat, pal):
    """The oracle's record of every ray of every hit pixel: (npx, nrays)."""
    rays = oracle.pixel_rays(vd, params, pos, nrm, mat, pal, np.arange(mat.size))
    return oracle.trace_rays(occ_h, rays.reshape(-1, 7), 0).reshape(mat.size, -1)


def compare_hits(got, ref, mat, what):
    sky = mat == 0
    assert not got[sky].view(np.uint8).any(), f"{what}: sky records not zero"
    g, r = got[~sky], ref[~sky]
    for f in ("term", "steps", "fetches", "voxel", "t", "normal"):
        assert_bitwise(g[f], r[f], f"{what}.{f}")
    return r


def _scene(api, ctx, oracle, model_dims, vol_dims, seed):
    model = api.Model(ctx, *model_dims)
    model.generate(seed=seed, terrain_fill=0.25, nboxes=30, box_min=2, box_max=12)
    model.build_mips()
    vol = model.download(0)
    occ = api.Volume(ctx, *vol_dims, nlevels=1)
    occ.clear()
    occ.rebuild([api.make_instance(model.mip_ptr(0), model.dims, IDENTITY, IDENTITY)])
    occ.build_levels()
    occ_h = np.zeros(vol_dims[::-1], np.uint8)
    oracle.rebuild(occ_h, [oracle.make_instance(vol, IDENTITY, IDENTITY)])
    return model, occ, occ_h


@pytest.mark.parametrize("vol_dims", [(45, 37, 43), (48, 40, 44)])
def test_hits_match_oracle_every_tuning(ctx, oracle, vol_dims):
    from voxelengine_amd import api, configs, views
    model, occ, occ_h = _scene(api, ctx, oracle, (96, 80, 88), vol_dims, seed=17)
    params = configs.trace_params(n_shadow=1, n_ao=6, ao_maxt=30.0, n_spec=1, march="dda")
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    w, h = 80, 56
    try:
        for tu in TUNINGS:
            ctx.set_tuning(**tu)
            for variant in (0, 1, 2, 3):
                vd = views.scene_view(model.dims, w, h, frame=variant, variant=variant)
                g = api.GBuffer(w, h)
                api.primary(ctx, model, vd, g)
                vis = api.Visibility(w, h, fetches=True)
                got = api.trace_hits(ctx, occ, vd, g, pal_d, params, vis)
                pos, nrm, mat = g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy()
                ref = oracle_hits(oracle, occ_h, vd, params, pos, nrm, mat, pal)
                r = compare_hits(got, ref, mat, f"{vol_dims} {tu} view {variant}")
                # the visibility written beside the records is the plain trace's
                sh, ao, sp, fe = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal)
                assert_bitwise(vis.ao.cpu().numpy(), ao, "ao")
                assert_bitwise(vis.fetches.cpu().numpy().astype(np.uint32), fe, "fetches")
                terms = set(np.unique(r["term"]).tolist())
                assert {1, 2} <= terms, terms        # hits and exits through the faces
    finally:
        ctx.reset_tuning()


@pytest.mark.parametrize("direction", [(1, 1, 1), (-1, 1, -1), (1, 1, 2), (2, -1, -1)])
def test_hits_on_exact_ties(ctx, oracle, direction):
    """Shadow rays from voxel centres along directions with equal components:
    two or three t_max tie at every step, so the hit face has two or three
    selected axes and best_t is the doubled or tripled t_max."""
    from voxelengine_amd import api, configs, views
    rng = np.random.default_rng(43)
    dims = (40, 40, 40)
    occ_h = ((rng.random(dims[::-1]) < 0.04).astype(np.uint8) *
             rng.integers(1, 256, dims[::-1]).astype(np.uint8))
    occ = api.Volume(ctx, *dims, nlevels=1)
    occ.clear()
    occ.upload_region(occ_h, [(0, 0, 0, *dims)])
    w, h = 32, 32
    vd = views.scene_view((80, 80, 80), w, h)
    g = api.GBuffer(w, h)
    cells = rng.integers(2, 76, (w * h, 3)).astype(np.float32) + 0.5
    g.pos.copy_(torch.from_numpy(cells))
    g.normal.zero_()
    g.material.fill_(20)
    d = np.asarray(direction, np.float32)
    d = (d / np.float32(np.sqrt(np.float32((d * d).sum())))).astype(np.float32)
    pal = configs.make_palette()
    params = configs.trace_params(n_shadow=1, n_ao=0, n_spec=0, shadow_maxt=128.0, march="dda", jitter_shadow=0)
    params["light_dir"] = [float(x) for x in d]
    got = api.trace_hits(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, api.Visibility(w, h))
    ref = oracle_hits(oracle, occ_h, vd, params, cells, np.zeros((w * h, 3), np.float32),
                      np.full(w * h, 20, np.uint8), pal)
    r = compare_hits(got, ref, np.full(w * h, 20, np.uint8), f"ties {direction}")
    hit = r["term"][:, 0] == 1
    assert hit.any()
    nsel = (r["normal"][hit, 0] != 0).sum(1)
    if len(set(abs(c) for c in direction)) == 1:
        assert (nsel >= 2).all()              # all components equal: every hit face is a tie
    else:
        assert (nsel >= 2).any()              # two equal components: some hit faces are ties


def test_hits_at_the_step_limit_below_a_ceiling(ctx, oracle):
    """Spec rays from a floor at fine y in [3, 4) toward an occupied ceiling at
    y = 280: an upward ray needs at least 276 steps to reach it, so the 256-step
    limit ends it first (term 0, 4 x 256 steps with the x4 restart) and it does
    not see the ceiling; the spec value stays 1. Every tuning, with and
    without the step-counting kernel."""
    from voxelengine_amd import api, configs, views
    S = (400, 320, 400)                       # wide enough that 256 steps cannot leave through a side
    host = np.zeros(S[::-1], np.uint8)
    host[:, :1, :] = 9                        # floor: fine y < 2
    host[:, 140:145, :] = 9                   # ceiling: fine y 280..289
    occ = api.Volume(ctx, *S, nlevels=1)
    occ.upload_region(host, [(0, 0, 0, *S)])
    w, h = 64, 48
    rng = np.random.default_rng(5)
    pos = np.c_[rng.uniform(300, 500, w * h), np.full(w * h, 2.0), rng.uniform(300, 500, w * h)].astype(np.float32)
    g = api.GBuffer(w, h)
    g.pos.copy_(torch.from_numpy(pos))
    g.normal.copy_(torch.tensor([0.0, 1.0, 0.0]).expand(w * h, 3))
    g.material.fill_(7)
    pal = configs.make_palette()
    pal[7, 3] = 1.0                           # roughness 1: directions over the whole hemisphere
    pal_d = torch.from_numpy(pal).cuda()
    vd = views.make_view((40.0, 40.0, 40.0), (40.0, 0.0, 41.0), w, h)
    params = configs.trace_params(n_shadow=0, n_ao=0, n_spec=1, spec_maxt=2000.0, march="dda")
    nrm, mat = np.tile(np.float32([0, 1, 0]), (w * h, 1)), np.full(w * h, 7, np.uint8)
    ref = oracle_hits(oracle, host, vd, params, pos, nrm, mat, pal)
    _, _, sp_ref, fe_ref = oracle.trace(host, vd, params, pos, nrm, mat, pal)
    rays = oracle.pixel_rays(vd, params, pos, nrm, mat, pal, np.arange(w * h))
    up = rays[:, 0, 4] > 0.05
    assert up.sum() > 0.3 * up.size
    assert (ref["term"][up, 0] == 0).all() and (ref["steps"][up, 0] == 1024).all()
    assert (sp_ref[up] == 1.0).all()
    try:
        for tu in TUNINGS:
            ctx.set_tuning(**tu)
            got = api.trace_hits(ctx, occ, vd, g, pal_d, params, api.Visibility(w, h))
            compare_hits(got, ref, mat, f"ceiling {tu}")
            vis = api.Visibility(w, h)
            api.trace(ctx, occ, vd, g, pal_d, params, vis)
            assert_bitwise(vis.spec.cpu().numpy(), sp_ref, f"spec without counts {tu}")
    finally:
        ctx.reset_tuning()
