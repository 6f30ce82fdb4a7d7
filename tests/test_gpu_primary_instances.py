"""Multi-instance geometry pass on the device (k_bin_instances + k_primary_inst)
against the oracle, bit for bit: random rotated / scaled / translated draws,
draws behind and around the camera, a tile with more draws than its list
holds, and the FarmHouse prefab's own draws feeding the trace."""
import numpy as np
import pytest

from helpers import assert_bitwise
from voxelengine_amd import views

pytestmark = pytest.mark.gpu


def _model(seed, dims, fill=0.08, floor=True):
    rng = np.random.default_rng(seed)
    sx, sy, sz = dims
    m = (rng.random((sz, sy, sx)) < fill).astype(np.uint8) * rng.integers(1, 255, (sz, sy, sx), dtype=np.uint8)
    if floor:
        m[:, :2, :] = rng.integers(1, 255, dtype=np.uint8)      # a floor in y
    return m


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _world(R, s, t):
    M = np.eye(4)
    M[:3, :3] = R * s
    M[:3, 3] = t
    return np.ascontiguousarray(M.T.reshape(16), np.float32)     # column-major


def _run(ctx, oracle, models_np, draws, vd):
    """draws: (model index, world, last_world, rid, palette)."""
    from voxelengine_amd import api
    dev = []
    for m in models_np:
        sz, sy, sx = m.shape
        mm = api.Model(ctx, sx, sy, sz)
        mm.upload(m)
        mm.build_mips()
        dev.append(mm)
    cmds = [api.make_vox_cmd(dev[k], w, lw, rid, pal) for (k, w, lw, rid, pal) in draws]
    W, H = vd["width"], vd["height"]
    g = api.GBuffer(W, H)
    ext = api.GBufferExt(W, H)
    api.primary_instances(ctx, cmds, vd, g, ext)
    import torch
    torch.cuda.synchronize()
    mips = [oracle.build_mips(m) for m in models_np]
    ref = oracle.primary_instances([dict(mips=mips[k], world=w, last_world=lw, rid=rid, palette=pal)
                                    for (k, w, lw, rid, pal) in draws], vd)
    got = dict(pos=g.pos.cpu().numpy(), nrm=g.normal.cpu().numpy(), mat=g.material.cpu().numpy(),
               motion=ext.motion.cpu().numpy(), depth=ext.depth.cpu().numpy(),
               palette=ext.palette.cpu().numpy().view(np.uint16))
    for k in ("pos", "nrm", "mat", "motion", "depth", "palette"):
        assert_bitwise(got[k].reshape(ref[k].shape), ref[k], k)
    return got, dev


def test_random_draws_match_oracle(ctx, oracle):
    rng = np.random.default_rng(11)
    models = [_model(1, (32, 32, 32)), _model(2, (16, 24, 40)), _model(3, (48, 8, 20), fill=0.2),
              _model(9, (32, 32, 32), fill=0.002, floor=False)]
    draws = []
    for k in range(10):
        mi = k % 3
        w = _world(_rot(rng), rng.uniform(0.5, 1.6), rng.uniform(-3.0, 6.0, 3))
        lw = _world(_rot(rng), rng.uniform(0.5, 1.6), rng.uniform(-3.0, 6.0, 3))
        draws.append((mi, w, lw, int(rng.integers(0, 5000)), k + 1))
    # one draw around the camera, one behind it
    draws.append((3, _world(np.eye(3), 1.0, (-1.6, 3.2, -1.6)), None, 7, 20))
    draws.append((1, _world(np.eye(3), 1.0, (-9.0, 6.0, -9.0)), None, 8, 21))
    draws = [(k, w, w if lw is None else lw, rid, p) for (k, w, lw, rid, p) in draws]
    vd = views.make_view((-0.5, 4.0, -0.5), (2.5, 0.5, 2.5), 160, 120, frame=3, jitter=(0.25, -0.375))
    vd_last = views.make_view((-0.7, 4.2, -0.4), (2.5, 0.5, 2.5), 160, 120)
    vd["last_view"] = vd_last["view"]
    got, _ = _run(ctx, oracle, models, draws, vd)
    assert (got["mat"] > 0).mean() > 0.2
    assert len(np.unique(got["palette"])) > 5


def test_tile_with_more_draws_than_its_list(ctx, oracle):
    """80 small draws in one screen region: tiles there take every draw."""
    rng = np.random.default_rng(12)
    models = [_model(4, (8, 8, 8), fill=0.5)]
    draws = []
    for k in range(80):
        t = np.array([2.0, 1.0, 2.0]) + rng.uniform(-0.4, 0.4, 3)
        w = _world(_rot(rng), 1.0, t)
        draws.append((0, w, w, k, k + 1))
    vd = views.make_view((0.0, 3.0, 0.0), (2.0, 1.0, 2.0), 96, 64)
    got, _ = _run(ctx, oracle, models, draws, vd)
    assert (got["mat"] > 0).sum() > 100


def test_farmhouse_draws_feed_the_trace(ctx, oracle, tmp_path):
    """FarmHouse spawned by the loader: its draws through the geometry pass,
    its instances through the rebuild, then shadow + AO rays over the
    instance G-buffer; every stage bit-exact against the oracle."""
    import torch
    from golden_mods import GOLDEN, write_mod_folder
    from voxelengine_amd import api, configs
    z = np.load(GOLDEN)
    mods = write_mod_folder(str(tmp_path))
    P = api.Prefab(mods, "default/FarmHouse.pf")
    models = P.load_models(ctx)
    cmds = P.vox_cmds(models)
    assert len(cmds) == 67
    files = [str(f) for f in z["FarmHouse_model_files"]]
    host = {fi: z[f"FarmHouse_model{files.index(P.files[fi].rsplit('/', 1)[1])}"] for fi in models}
    mips = {fi: oracle.build_mips(a) for fi, a in host.items()}
    by_handle = {m.h.value: fi for fi, m in models.items()}
    W, H = 240, 136
    vd = views.make_view((8.0, 16.0, 6.0), (32.0, 0.0, 32.0), W, H, frame=1)
    g = api.GBuffer(W, H)
    ext = api.GBufferExt(W, H)
    api.primary_instances(ctx, cmds, vd, g, ext)
    torch.cuda.synchronize()
    ref = oracle.primary_instances([dict(mips=mips[by_handle[c.model]], world=np.array(c.world),
                                         last_world=np.array(c.last_world), rid=c.volume_rid, palette=c.palette)
                                    for c in cmds], vd)
    assert_bitwise(g.pos.cpu().numpy(), ref["pos"], "pos")
    assert_bitwise(g.normal.cpu().numpy(), ref["nrm"], "nrm")
    assert_bitwise(g.material.cpu().numpy(), ref["mat"], "mat")
    assert_bitwise(ext.depth.cpu().numpy(), ref["depth"], "depth")
    assert (ref["mat"] > 0).mean() > 0.3
    # the scene's occupancy, then shadow + AO over the instance G-buffer
    occ = api.Volume(ctx, 524, 188, 524, nlevels=2)
    occ.clear()
    occ.rebuild(P.instances(models))
    occ.build_levels()
    occ_h = occ.download()
    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    params = configs.trace_params(n_ao=4, ao_maxt=16.0, march="dda")
    vis = api.Visibility(W, H)
    api.trace(ctx, occ, vd, g, pal_d, params, vis)
    torch.cuda.synchronize()
    sh, ao, sp, _ = oracle.trace(occ_h, vd, params, ref["pos"], ref["nrm"], ref["mat"], pal)
    assert_bitwise(vis.shadow.cpu().numpy(), sh, "shadow")
    assert_bitwise(vis.ao.cpu().numpy(), ao, "ao")
    assert 0.05 < (sh[ref["mat"] > 0] > 0).mean() < 0.98


def test_geometry_pass_error_paths(ctx):
    """Limits and bad input come back as errors before any launch."""
    import torch
    from voxelengine_amd import api, _lib as L
    m = api.Model(ctx, 8, 8, 8)
    g = api.GBuffer(16, 16)
    vd = views.make_view((0.0, 2.0, 0.0), (1.0, 0.0, 1.0), 16, 16)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    with pytest.raises(L.VXError, match="at most 4096"):
        api.primary_instances(ctx, [api.make_vox_cmd(m, eye)] * 4097, vd, g)
    with pytest.raises(L.VXError, match="singular"):
        api.primary_instances(ctx, [api.make_vox_cmd(m, np.zeros(16, np.float32))], vd, g)
    with pytest.raises(L.VXError, match="palette"):
        api.primary_instances(ctx, [api.make_vox_cmd(m, eye, palette=70000)], vd, g)
    api.primary_instances(ctx, [], vd, g)           # no draws: every pixel misses
    torch.cuda.synchronize()
    assert int(g.material.sum().item()) == 0


def test_taa_error_paths(ctx):
    import torch
    from voxelengine_amd import api, configs, _lib as L
    W, H = 16, 16
    vd = views.make_view((0.0, 2.0, 0.0), (1.0, 0.0, 1.0), W, H)
    z = lambda *s, dt=torch.float32: torch.zeros(s, dtype=dt, device="cuda")
    pal = torch.from_numpy(configs.make_palette()).cuda()
    last = z(W * H, 4)
    with pytest.raises(L.VXError, match="alias"):
        api.light_taa(ctx, vd, z(W * H, 3), last, last, z(W * H), z(W * H, 3), z(W * H, dt=torch.uint8), pal)
    with pytest.raises(L.VXError, match="aligned"):
        api.light_taa(ctx, vd, z(W * H, 3), z(W * H * 4 + 1)[1:], z(W * H, 4), z(W * H), z(W * H, 3),
                      z(W * H, dt=torch.uint8), pal)
