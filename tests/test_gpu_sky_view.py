"""A view that sees no voxel at all (the camera above a corner of the scene,
looking up and away from it): every pixel is sky, so the ray-parallel trace,
its batched form, the light pass on both paths, the reflection pass and the
screen-space occlusion get no work. Each must still end and write the sky
values, bit for bit against the oracle, and a batch holding the sky view next
to a view of the scene must give both items the bits of their own launches.
The primary pass's coverage (which pixels the model's box is drawn at) is
also checked against the oracle for cameras beside, inside and below the
box."""
import numpy as np
import pytest
import torch

from helpers import assert_bitwise, device_scene, oracle_occ

pytestmark = pytest.mark.gpu

W, H = 96, 64


@pytest.fixture(scope="module")
def scene(ctx, oracle):
    from voxelengine_amd import api, views
    model, occ, vol = device_scene(api, ctx, 64, seed=29, fill=0.3, nboxes=25)
    nx, ny, nz = model.dims
    eye = np.array([-0.08 * nx, 1.05 * ny, -0.08 * nz]) / 10.0
    sky_vd = views.make_view(eye, eye + np.array([-1.0, 0.6, -1.0]), W, H, frame=1)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, sky_vd, g)
    torch.cuda.synchronize()
    return dict(model=model, occ=occ, vol=vol, occ_h=oracle_occ(oracle, vol), mips=oracle.build_mips(vol), vd=sky_vd,
                g=g)


def test_sky_view_has_no_hit(scene, oracle):
    g = scene["g"]
    mat = g.material.cpu().numpy()
    assert not mat.any()
    pos, nrm, omat = oracle.primary(scene["mips"], scene["vd"])
    assert not omat.any()
    assert_bitwise(g.pos.cpu().numpy(), pos, "pos")
    assert_bitwise(g.normal.cpu().numpy(), nrm, "normal")


def test_trace_and_hits_of_a_sky_view(ctx, oracle, scene):
    from voxelengine_amd import api, configs
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    params = configs.trace_params(n_shadow=1, n_ao=4, ao_maxt=20.0, n_spec=1, march="dda")
    g = scene["g"]
    vis = api.Visibility(W, H, fetches=True)
    api.trace(ctx, scene["occ"], scene["vd"], g, pal_d, params, vis)
    torch.cuda.synchronize()
    sh, ao, sp, fe = oracle.trace(scene["occ_h"], scene["vd"], params, g.pos.cpu().numpy(), g.normal.cpu().numpy(),
                                  g.material.cpu().numpy(), pal)
    assert_bitwise(vis.shadow.cpu().numpy(), sh, "shadow")
    assert_bitwise(vis.ao.cpu().numpy(), ao, "ao")
    assert_bitwise(vis.spec.cpu().numpy(), sp, "spec")
    assert_bitwise(vis.fetches.cpu().numpy().astype(np.uint32), fe, "fetches")
    assert not fe.any()
    # the step-counting launch: no ray, so every record stays zero
    hv = api.Visibility(W, H, fetches=True)
    hits = api.trace_hits(ctx, scene["occ"], scene["vd"], g, pal_d, params, hv)
    torch.cuda.synchronize()
    assert not hits.view(np.uint8).any()
    assert torch.equal(hv.ao.view(torch.int32), vis.ao.view(torch.int32))


def test_batch_with_a_sky_item(ctx, scene):
    from voxelengine_amd import api, configs, views
    pal_d = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_shadow=1, n_ao=4, ao_maxt=20.0, n_spec=1, march="dda")
    vd2 = views.scene_view(scene["model"].dims, 40, 24, frame=1)
    g2 = api.GBuffer(40, 24)
    api.primary(ctx, scene["model"], vd2, g2)
    plain = []
    for vd, g, w, h in [(scene["vd"], scene["g"], W, H), (vd2, g2, 40, 24)]:
        p = api.Visibility(w, h, fetches=True)
        api.trace(ctx, scene["occ"], vd, g, pal_d, params, p)
        plain.append(p)
    outs = [api.Visibility(W, H, fetches=True), api.Visibility(40, 24, fetches=True)]
    b = api.Batch(ctx, [(scene["occ"], scene["vd"], scene["g"], outs[0]), (scene["occ"], vd2, g2, outs[1])],
                  pal_d, params)
    try:
        # the sky item alone, then both
        for items in ([(scene["occ"], scene["vd"], scene["g"], outs[0])],
                      [(scene["occ"], scene["vd"], scene["g"], outs[0]), (scene["occ"], vd2, g2, outs[1])]):
            b.update(items)
            for _ in range(2):
                b.launch()
                torch.cuda.synchronize()
                for (_, _, _, o), p in zip(items, plain):
                    for f in ("shadow", "ao", "spec", "fetches"):
                        assert torch.equal(getattr(o, f).view(torch.uint8), getattr(p, f).view(torch.uint8)), f
    finally:
        b.close()
    assert (g2.material.cpu().numpy() > 0).any()


def test_light_reflection_and_occlusion_passes_of_a_sky_view(ctx, oracle, scene):
    from voxelengine_amd import api, configs
    vd, g = scene["vd"], scene["g"]
    npx = W * H
    pos, nrm, mat = g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy()
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    depth = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.gbuf_depth(ctx, vd, g, depth, volume_rid=1)
    s = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.ssao(ctx, vd, g, depth, s, seed=5)
    torch.cuda.synchronize()
    d = depth.cpu().numpy()
    assert (d == 1.0).all()
    assert_bitwise(s.cpu().numpy(), oracle.ssao(vd, nrm, mat, d, seed=5), "ssao")

    rng = np.random.default_rng(7)
    ext = np.array(scene["model"].dims, np.float64) * 0.1
    L = np.zeros(8, api.LIGHT_DTYPE)
    for k in range(L.size):
        L[k]["type"] = k % 2
        L[k]["position"] = rng.uniform([0.0, 0.3, 0.0], [1.0, 0.8, 1.0]) * ext
        L[k]["range"] = ext.max()
        L[k]["color"] = rng.uniform(0.5, 3.0, 3)
        L[k]["attenuation"] = 1.0
        dd = rng.normal(size=3)
        L[k]["direction"] = dd / np.linalg.norm(dd)
        L[k]["angle"] = 0.6
        L[k]["angle_attenuation"] = 1.0
    base = rng.random((npx, 3)).astype(np.float32)
    ref, ref_f = oracle.light_pass(scene["occ_h"], vd, L, oracle.light_params(1, 2, 11, 1), pos, nrm, mat, pal,
                                   rgb=base)
    assert_bitwise(ref, base, "oracle light pass of a sky view")
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

    light = rng.uniform(0, 3, (npx, 4)).astype(np.float32)
    out = torch.full((npx, 4), -1.0, dtype=torch.float32, device="cuda")
    t = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.reflection(ctx, scene["occ"], vd, g, pal_d, torch.from_numpy(light).cuda(), depth, out, t=t, march=1,
                   maxt=256.0, seed=2, frame=1)
    torch.cuda.synchronize()
    rref, tref = oracle.reflection(scene["occ_h"], vd, oracle.reflection_params(1, 256.0, 2, 1), pos, nrm, mat, pal,
                                   d, light)
    assert_bitwise(t.cpu().numpy(), tref, "reflection t")
    assert_bitwise(out.cpu().numpy(), rref, "reflection out")


CAMERAS = {
    "beside, half behind": (lambda n: (0.5 * n[0], 0.6 * n[1], -0.3), lambda n: (-2.0, 0.6 * n[1], 0.1)),
    "inside": (lambda n: (0.5 * n[0], 0.5 * n[1], 0.5 * n[2]), lambda n: (0.9 * n[0], 0.3 * n[1], 0.2 * n[2])),
    "below, looking up": (lambda n: (0.4 * n[0], -0.5, 0.6 * n[2]), lambda n: (0.5 * n[0], 2.0 * n[1], 0.5 * n[2])),
    "corner, facing the scene": (lambda n: (-0.1 * n[0], 1.1 * n[1], -0.1 * n[2]),
                                 lambda n: (0.5 * n[0], 0.2 * n[1], 0.5 * n[2])),
}


def _camera(scene, camera):
    """Eye and target (world units) of a named camera. The inside camera sits
    in the first candidate point whose voxel and 4-voxel neighbourhood are
    empty (a camera inside a set voxel hits it at t = 0 everywhere, and the
    passes' view vectors are then NaN)."""
    n = np.array(scene["model"].dims, np.float64) / 10.0
    eye, tgt = CAMERAS[camera]
    e = np.array(eye(n))
    if camera == "inside":
        vol = scene["vol"]                                  # (z, y, x) fine voxels
        for c in ([0.5, 0.5, 0.5], [0.5, 0.85, 0.5], [0.3, 0.9, 0.7], [0.7, 0.95, 0.3], [0.5, 0.97, 0.5]):
            v = (np.array(c) * np.array(scene["model"].dims)).astype(int)
            lo, hi = np.maximum(v - 4, 0), v + 5
            if not vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]].any():
                e = (v + 0.5) / 10.0
                break
        else:
            raise AssertionError("no empty point for the inside camera")
    return e, np.array(tgt(n))


@pytest.mark.parametrize("camera", list(CAMERAS), ids=[c.replace(", ", "-").replace(" ", "_") for c in CAMERAS])
def test_primary_coverage_matches_oracle(ctx, oracle, scene, camera):
    """Where the model's box is drawn decides which pixels are marched: the
    kernel and the oracle agree bit for bit for cameras beside the box (part
    of it behind the camera plane), inside it, below it and facing it."""
    from voxelengine_amd import api, views
    eye, tgt = _camera(scene, camera)
    vd = views.make_view(eye, tgt, W, H, frame=1)
    g = api.GBuffer(W, H)
    api.primary(ctx, scene["model"], vd, g)
    torch.cuda.synchronize()
    pos, nrm, mat = oracle.primary(scene["mips"], vd)
    assert_bitwise(g.material.cpu().numpy(), mat, f"{camera}: material")
    assert_bitwise(g.pos.cpu().numpy(), pos, f"{camera}: pos")
    assert_bitwise(g.normal.cpu().numpy(), nrm, f"{camera}: normal")
    assert (mat > 0).any()


@pytest.mark.parametrize("camera", ["beside, half behind", "inside", "below, looking up"],
                         ids=["beside-half_behind", "inside", "below-looking_up"])
def test_passes_of_other_cameras_match_oracle(ctx, oracle, scene, camera):
    """The passes after the primary hit for the same cameras: the D24 depth
    (its range), screen-space occlusion, the light pass and the reflection
    pass (whose reprojection and sky depend on where the camera sits), bit
    for bit against the oracle."""
    from voxelengine_amd import api, configs, views
    n = np.array(scene["model"].dims, np.float64) / 10.0
    eye, tgt = _camera(scene, camera)
    vd = views.make_view(eye, tgt, W, H, frame=1)
    npx = W * H
    g = api.GBuffer(W, H)
    api.primary(ctx, scene["model"], vd, g)
    depth = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.gbuf_depth(ctx, vd, g, depth, volume_rid=1)
    s = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.ssao(ctx, vd, g, depth, s, seed=3)
    torch.cuda.synchronize()
    pos, nrm, mat = g.pos.cpu().numpy(), g.normal.cpu().numpy(), g.material.cpu().numpy()
    d = depth.cpu().numpy()
    assert ((d >= 0) & (d <= 1)).all()
    assert_bitwise(s.cpu().numpy(), oracle.ssao(vd, nrm, mat, d, seed=3), f"{camera}: ssao")

    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    rng = np.random.default_rng(13)
    ext = n
    L = np.zeros(6, api.LIGHT_DTYPE)
    for k in range(L.size):
        L[k]["type"] = k % 2
        L[k]["position"] = rng.uniform([0.1, 0.2, 0.1], [0.9, 0.8, 0.9]) * ext
        L[k]["range"] = 0.6 * ext.max()
        L[k]["color"] = rng.uniform(0.5, 3.0, 3)
        L[k]["attenuation"] = 1.5
        dd = rng.normal(size=3)
        L[k]["direction"] = dd / np.linalg.norm(dd)
        L[k]["angle"] = 0.7
        L[k]["angle_attenuation"] = 1.0
    base = rng.random((npx, 3)).astype(np.float32)
    ref, ref_f = oracle.light_pass(scene["occ_h"], vd, L, oracle.light_params(1, 2, 11, 1), pos, nrm, mat, pal,
                                   rgb=base)
    rgb = torch.from_numpy(base.copy()).cuda()
    fetches = torch.full((npx,), -1, dtype=torch.int32, device="cuda")
    api.light_pass(ctx, scene["occ"], vd, g, pal_d, api.lights_to_device(L), L.size, rgb, point_march=1,
                   spot_march=2, seed=11, frame=1, fetches=fetches)
    light = rng.uniform(0, 3, (npx, 4)).astype(np.float32)
    out = torch.full((npx, 4), -1.0, dtype=torch.float32, device="cuda")
    t = torch.full((npx,), -1.0, dtype=torch.float32, device="cuda")
    api.reflection(ctx, scene["occ"], vd, g, pal_d, torch.from_numpy(light).cuda(), depth, out, t=t, march=0,
                   maxt=256.0, seed=2, frame=1)
    torch.cuda.synchronize()
    assert_bitwise(rgb.cpu().numpy(), ref, f"{camera}: light rgb")
    assert_bitwise(fetches.cpu().numpy().astype(np.uint32), ref_f, f"{camera}: light fetches")
    rref, tref = oracle.reflection(scene["occ_h"], vd, oracle.reflection_params(0, 256.0, 2, 1), pos, nrm, mat, pal,
                                   d, light)
    assert_bitwise(t.cpu().numpy(), tref, f"{camera}: reflection t")
    assert_bitwise(out.cpu().numpy(), rref, f"{camera}: reflection out")
