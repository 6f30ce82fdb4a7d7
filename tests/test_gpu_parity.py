"""GPU parity: every kernel of the path against the oracle on the same inputs.

Integer / index outputs and the fp32 traversal are compared bit for bit (the
kernels and the oracle evaluate the same IEEE fp32 expressions without FMA
contraction). Gradients: 1e-4 relative to the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from helpers import (IDENTITY, assert_bitwise, compare_ray_results, device_scene, edge_rays, oracle_occ,
                     random_rays)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from voxelengine_amd import api as A
    return A


@pytest.fixture(scope="module")
def scene64(api, ctx):
    return device_scene(api, ctx, 64, seed=5, fill=0.3, nboxes=40, box=(2, 16), levels=2)


@pytest.mark.parametrize("dims,fill,nboxes,grounded", [((64, 64, 64), 0.3, 40, False),
                                                       ((72, 40, 96), 0.15, 25, True),
                                                       ((128, 64, 128), 0.0, 300, True)])
def test_scene_generation_matches_oracle(api, ctx, oracle, dims, fill, nboxes, grounded):
    m = api.Model(ctx, *dims)
    m.generate(seed=77, terrain_fill=fill, nboxes=nboxes, box_min=2, box_max=20, grounded=grounded)
    got = m.download(0)
    ref = oracle.gen_terrain(*m.dims, 77, fill) if fill > 0 else np.zeros(m.dims[::-1], np.uint8)
    oracle.gen_boxes(ref, 77, nboxes, 2, 20, int(grounded))
    assert_bitwise(got, ref, "scene")


def test_mips_match_oracle(api, scene64, oracle):
    model, _, vol = scene64
    ref = oracle.build_mips(vol)
    for k in (1, 2):
        assert_bitwise(model.download(k), ref[k], f"mip{k}")


def test_rebuild_identity_matches_oracle(api, scene64, oracle):
    _, occ, vol = scene64
    assert_bitwise(occ.download(), oracle_occ(oracle, vol), "occupancy")


def test_levels_are_or_reductions(api, scene64, oracle):
    _, occ, vol = scene64
    bytes_ = oracle_occ(oracle, vol)
    S = bytes_.shape[::-1]
    # level 1: one bit per 4^3 texels (8^3 fine voxels)
    B = tuple(s // 4 + 1 for s in S)
    pad = np.zeros((B[2] * 4, B[1] * 4, B[0] * 4), np.uint8)
    pad[:S[2], :S[1], :S[0]] = bytes_
    l1 = pad.reshape(B[2], 4, B[1], 4, B[0], 4).any(axis=(1, 3, 5))
    words = occ.download_level(1)
    W = tuple((b + 3) // 4 for b in B)
    got = np.zeros((W[2] * 4, W[1] * 4, W[0] * 4), bool)
    for wz in range(W[2]):
        for wy in range(W[1]):
            for wx in range(W[0]):
                w = int(words[(wz * W[1] + wy) * W[0] + wx])
                for b in range(64):
                    got[wz * 4 + (b >> 4), wy * 4 + ((b >> 2) & 3), wx * 4 + (b & 3)] = (w >> b) & 1
    assert np.array_equal(got[:B[2], :B[1], :B[0]], l1)
    assert not got[B[2]:].any() and not got[:, B[1]:].any() and not got[:, :, B[0]:].any()
    # level 2: one bit per 8^3 level-1 cells
    C2 = tuple((b + 7) // 8 for b in B)
    p1 = np.zeros((C2[2] * 8, C2[1] * 8, C2[0] * 8), bool)
    p1[:B[2], :B[1], :B[0]] = l1
    l2 = p1.reshape(C2[2], 8, C2[1], 8, C2[0], 8).any(axis=(1, 3, 5))
    w2 = occ.download_level(2)
    W2 = tuple((c + 3) // 4 for c in C2)
    for cz in range(C2[2]):
        for cy in range(C2[1]):
            for cx in range(C2[0]):
                w = int(w2[((cz >> 2) * W2[1] + (cy >> 2)) * W2[0] + (cx >> 2)])
                b = (cx & 3) | ((cy & 3) << 2) | ((cz & 3) << 4)
                assert bool((w >> b) & 1) == bool(l2[cz, cy, cx])


def _rot(ax, ang):
    c, s = np.cos(ang), np.sin(ang)
    m = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][ax]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def _colmajor(m):
    return np.ascontiguousarray(np.asarray(m, np.float64).T.reshape(16), dtype=np.float32)


def test_rebuild_instances_match_oracle(api, ctx, oracle):
    """Translated, rotated, scaled and pivoted instances, overlapping footprints,
    erase-before-write ordering, a destroyed instance and an invalid (NULL) model."""
    rng = np.random.default_rng(3)
    S = (48, 40, 56)
    models = []
    for k in range(5):
        dims = tuple(int(4 * rng.integers(2, 9)) for _ in range(3))
        mvol = (rng.integers(0, 40, size=dims[::-1]) * (rng.random(dims[::-1]) < 0.4)).astype(np.uint8)
        models.append(mvol)
    host = np.zeros(S[::-1], np.uint8)
    occ = api.Volume(ctx, *S, nlevels=2)
    occ.clear()
    dev = [torch.from_numpy(m.reshape(-1)).cuda() for m in models]
    inst_o, inst_d = [], []
    for k, mvol in enumerate(models):
        T = np.eye(4)
        T[:3, 3] = rng.uniform(0.0, 6.0, 3)
        M = T @ _rot(k % 3, rng.uniform(-1.0, 1.0)) @ np.diag([1.0, rng.uniform(0.7, 1.3), 1.0, 1.0])
        Mp = np.eye(4)
        Mp[:3, 3] = rng.uniform(0.0, 6.0, 3)
        prev, cur = _colmajor(Mp), _colmajor(M)
        piv = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
        flags = 2 if k == 3 else 1
        dims = mvol.shape[::-1]
        inst_o.append(oracle.make_instance(mvol, prev, cur, piv, flags))
        inst_d.append(api.make_instance(dev[k].data_ptr(), dims, prev, cur, piv, flags))
    inst_o.append(oracle.make_instance(None, IDENTITY, IDENTITY))
    inst_d.append(api.make_instance(0, (4, 4, 4), IDENTITY, IDENTITY))
    # two frames so the erase of the previous footprint matters
    for frame in range(2):
        ro = oracle.rebuild(host, inst_o)
        rd = occ.rebuild(inst_d)
        assert [tuple(r) for r in ro.tolist()] == [tuple(r) for r in rd], (frame, ro, rd)
        assert_bitwise(occ.download(), host, f"occupancy frame {frame}")
        for io, idd in zip(inst_o, inst_d):
            io.prev = io.cur
            idd.prev = idd.cur


def test_rebuild_scaled_ragged_and_unaligned_instances_match_oracle(api, ctx, oracle):
    """A x8 scaled instance (its footprint per block of model voxels is too
    large for the shared-memory window, so its bits go out one atomic per
    voxel word), rows that are not whole quads, and a model that starts at an
    odd address; two frames so the erase matters."""
    rng = np.random.default_rng(21)
    S = (40, 36, 44)
    dims = [(13, 7, 9), (70, 5, 6), (16, 12, 20)]
    models = [(rng.integers(0, 40, size=d[::-1]) * (rng.random(d[::-1]) < 0.5)).astype(np.uint8) for d in dims]
    host = np.zeros(S[::-1], np.uint8)
    occ = api.Volume(ctx, *S, nlevels=1)
    occ.clear()
    # the third model sits one byte into its buffer
    raw = [torch.from_numpy(m.reshape(-1)).cuda() for m in models[:2]]
    buf = torch.zeros(models[2].size + 1, dtype=torch.uint8, device="cuda")
    buf[1:] = torch.from_numpy(models[2].reshape(-1)).cuda()
    ptrs = [raw[0].data_ptr(), raw[1].data_ptr(), buf.data_ptr() + 1]
    scales = [np.diag([8.0, 8.0, 8.0, 1.0]), _rot(1, 0.6), _rot(0, -0.3) @ np.diag([1.0, 1.2, 0.9, 1.0])]
    inst_o, inst_d = [], []
    for k, (mvol, M0) in enumerate(zip(models, scales)):
        T = np.eye(4)
        T[:3, 3] = rng.uniform(0.5, 2.0, 3)
        M = T @ M0
        Mp = np.eye(4)
        Mp[:3, 3] = rng.uniform(0.0, 3.0, 3)
        prev, cur = _colmajor(Mp @ M0), _colmajor(M)
        piv = rng.uniform(-0.5, 0.5, 3).astype(np.float32)
        inst_o.append(oracle.make_instance(mvol, prev, cur, piv, 1))
        inst_d.append(api.make_instance(ptrs[k], mvol.shape[::-1], prev, cur, piv, 1))
    for frame in range(2):
        ro = oracle.rebuild(host, inst_o)
        rd = occ.rebuild(inst_d)
        assert [tuple(r) for r in ro.tolist()] == [tuple(r) for r in rd], (frame, ro, rd)
        assert_bitwise(occ.download(), host, f"occupancy frame {frame}")
        assert host.any()
        for io, idd in zip(inst_o, inst_d):
            io.prev = io.cur
            idd.prev = idd.cur


def test_upload_download_roundtrip(api, ctx):
    rng = np.random.default_rng(9)
    S = (20, 12, 28)
    host = (rng.random(S[::-1]) < 0.3).astype(np.uint8) * rng.integers(1, 256, S[::-1]).astype(np.uint8)
    occ = api.Volume(ctx, *S, nlevels=1)
    occ.clear()
    occ.upload_region(host, [(0, 0, 0, S[0], S[1], S[2])])
    assert_bitwise(occ.download(), host, "roundtrip")
    host2 = host.copy()
    host2[3:7, 2:5, 4:9] = 0xA5
    occ.upload_region(host2, [(4, 2, 3, 5, 3, 4)])
    assert_bitwise(occ.download(), host2, "region upload")


@pytest.mark.parametrize("march", [0, 1, 2])
def test_trace_rays_random_and_edges(api, ctx, scene64, oracle, march):
    _, occ, vol = scene64
    occ_h = oracle_occ(oracle, vol)
    S = occ_h.shape[0]
    rng = np.random.default_rng(100 + march)
    rays = np.concatenate([random_rays(rng, 20000, -6.0, 2 * S + 6.0), edge_rays(S)])
    ref = oracle.trace_rays(occ_h, rays, march)
    got = api.trace_rays(ctx, occ, torch.from_numpy(rays).cuda(), march)
    compare_ray_results(got, ref, f"march{march}")
    assert ref["hit"].sum() > 100


def test_primary_matches_oracle(api, ctx, scene64, oracle):
    from voxelengine_amd import views
    model, _, vol = scene64
    mips = oracle.build_mips(vol)
    for variant, (w, h) in enumerate([(96, 64), (77, 45)]):
        vd = views.scene_view(model.dims, w, h, frame=variant, variant=variant)
        g = api.GBuffer(w, h)
        api.primary(ctx, model, vd, g)
        pos, nrm, mat = oracle.primary(mips, vd)
        assert_bitwise(g.material.cpu().numpy(), mat, "material")
        assert_bitwise(g.pos.cpu().numpy(), pos, "pos")
        assert_bitwise(g.normal.cpu().numpy(), nrm, "normal")
        assert (mat > 0).mean() > 0.3


@pytest.mark.parametrize("march", ["dda", "reference"])
def test_trace_pass_matches_oracle(api, ctx, scene64, oracle, march):
    from voxelengine_amd import configs, views
    model, occ, vol = scene64
    occ_h = oracle_occ(oracle, vol)
    w, h = 80, 56
    vd = views.scene_view(model.dims, w, h, variant=1)
    g = api.GBuffer(w, h)
    api.primary(ctx, model, vd, g)
    pal = configs.make_palette()
    params = configs.trace_params(n_ao=6, ao_maxt=24.0, march=march)
    vis = api.Visibility(w, h, fetches=True)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
    pos, nrm, mat = (g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy())
    sh, ao, sp, fe = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal)
    assert_bitwise(vis.shadow.cpu().numpy(), sh, "shadow")
    assert_bitwise(vis.ao.cpu().numpy(), ao, "ao")
    assert_bitwise(vis.spec.cpu().numpy(), sp, "spec")
    assert_bitwise(vis.fetches.cpu().numpy().astype(np.uint32), fe, "fetches")
    hit = mat > 0
    assert 0.0 < sh[hit].mean() < 1.0


@pytest.mark.parametrize("jitter", [0, 1])
def test_trace_pass_crafted_gbuffer(api, ctx, scene64, oracle, jitter):
    """A hand-made G-buffer with integral positions, zero and axis-aligned normals:
    rays outside the DDA kernel's fast path (zero or tiny direction components,
    origins on voxel boundaries) take the deferred exact path; results stay
    bit-identical to the oracle."""
    from voxelengine_amd import configs, views
    model, occ, vol = scene64
    occ_h = oracle_occ(oracle, vol)
    w, h = 40, 30
    n = w * h
    rng = np.random.default_rng(8)
    pos = rng.uniform(4.0, 60.0, size=(n, 3)).astype(np.float32)
    pos[::3] = np.round(pos[::3])
    nrm = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    nrm[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    nrm[::7] = 0.0
    nrm[1::11] = rng.normal(size=(len(nrm[1::11]), 3))
    mat = rng.integers(0, 256, n).astype(np.uint8)
    mat[::13] = 0
    vd = views.scene_view(model.dims, w, h)
    g = api.GBuffer(w, h)
    g.pos.copy_(torch.from_numpy(pos))
    g.normal.copy_(torch.from_numpy(nrm))
    g.material.copy_(torch.from_numpy(mat))
    pal = configs.make_palette()
    params = configs.trace_params(n_ao=5, ao_maxt=20.0, march="dda", jitter_shadow=jitter)
    vis = api.Visibility(w, h, fetches=True)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
    sh, ao, sp, fe = oracle.trace(occ_h, vd, params, pos, nrm, mat, pal)
    assert_bitwise(vis.shadow.cpu().numpy(), sh, "shadow")
    assert_bitwise(vis.ao.cpu().numpy(), ao, "ao")
    assert_bitwise(vis.spec.cpu().numpy(), sp, "spec")
    assert_bitwise(vis.fetches.cpu().numpy().astype(np.uint32), fe, "fetches")


def _shade_inputs(api, ctx, scene64, oracle, w=64, h=48):
    from voxelengine_amd import configs, views
    model, occ, vol = scene64
    vd = views.scene_view(model.dims, w, h, variant=2)
    g = api.GBuffer(w, h)
    api.primary(ctx, model, vd, g)
    pal = configs.make_palette()
    params = configs.trace_params(n_ao=4, ao_maxt=16.0, n_spec=0)
    vis = api.Visibility(w, h)
    api.trace(ctx, occ, vd, g, torch.from_numpy(pal).cuda(), params, vis)
    return vd, g, vis, pal


def test_shade_forward_matches_oracle(api, ctx, scene64, oracle):
    from voxelengine_amd import configs
    vd, g, vis, pal = _shade_inputs(api, ctx, scene64, oracle)
    light = configs.default_light()
    n = vd["width"] * vd["height"]
    rgb = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    api.shade_forward(ctx, vd, g, vis, torch.from_numpy(pal).cuda(), light, rgb)
    args = (g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(), vis.shadow.cpu().numpy(),
            vis.ao.cpu().numpy())
    ref = oracle.shade_f32(vd, *args, pal, light)
    assert_bitwise(rgb.cpu().numpy(), ref, "rgb")
    ref64 = oracle.shade_f64(vd, *args, pal.astype(np.float64), light.astype(np.float64))
    assert np.abs(ref64 - rgb.cpu().numpy()).max() < 1e-5


def test_shade_backward_matches_fp64(api, ctx, scene64, oracle):
    from voxelengine_amd import configs
    vd, g, vis, pal = _shade_inputs(api, ctx, scene64, oracle)
    light = configs.default_light()
    n = vd["width"] * vd["height"]
    rng = np.random.default_rng(4)
    dl = rng.normal(size=(n, 3)).astype(np.float32)
    dpal = torch.zeros((256, 6), dtype=torch.float32, device="cuda")
    dlight = torch.zeros(6, dtype=torch.float32, device="cuda")
    pal_d = torch.from_numpy(pal).cuda()
    api.shade_backward(ctx, vd, g, vis, pal_d, light, torch.from_numpy(dl).cuda(), dpal, dlight)
    args = (g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy(), vis.shadow.cpu().numpy(),
            vis.ao.cpu().numpy())
    rp, rl = oracle.shade_backward_f64(vd, *args, pal.astype(np.float64), light.astype(np.float64),
                                       dl.astype(np.float64))
    gp, gl = dpal.cpu().numpy(), dlight.cpu().numpy()
    scale_p = np.abs(rp).max()
    assert np.abs(gp - rp).max() <= 1e-4 * scale_p
    assert np.abs(gl - rl).max() <= 1e-4 * np.abs(rl).max()
    # deterministic: a second launch gives identical bits
    dpal2 = torch.zeros_like(dpal)
    dlight2 = torch.zeros_like(dlight)
    api.shade_backward(ctx, vd, g, vis, pal_d, light, torch.from_numpy(dl).cuda(), dpal2, dlight2)
    assert_bitwise(dpal2.cpu().numpy(), gp, "dpalette rerun")
    assert_bitwise(dlight2.cpu().numpy(), gl, "dlight rerun")
