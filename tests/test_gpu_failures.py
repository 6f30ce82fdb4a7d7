"""Error paths on the device, forced from the host: an injected device
allocation failure or kernel-launch failure makes the call report
VX_ERR_OUT_OF_MEMORY / VX_ERR_DEVICE and leaves the objects usable, so the
same call made again gives the bits of an undisturbed run; a batch whose
update failed refuses to launch until an update succeeds; a batch re-plans
and re-captures its launch when its items change shape. Every kernel runs to
its normal end: the failures are reported by the host-side checks."""
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

from helpers import IDENTITY, device_scene

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def disarm():
    from voxelengine_amd import api
    yield api
    api.inject_failure(api.FAIL_ALLOC, -1)
    api.inject_failure(api.FAIL_LAUNCH, -1)


@pytest.fixture(scope="module")
def scene(ctx):
    from voxelengine_amd import api, configs, views
    model, occ, vol = device_scene(api, ctx, 64, seed=12, nboxes=20, box=(3, 12))
    W, H = 96, 64
    vd = views.scene_view(model.dims, W, H, frame=2)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, vd, g)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_shadow=1, n_ao=4, ao_maxt=24.0, n_spec=1, march="dda")
    ref = api.Visibility(W, H, fetches=True)
    api.trace(ctx, occ, vd, g, pal, params, ref)
    torch.cuda.synchronize()
    return dict(model=model, occ=occ, vol=vol, vd=vd, g=g, pal=pal, params=params, ref=ref, W=W, H=H)


def _same(a, b):
    for f in ("shadow", "ao", "spec", "fetches"):
        assert torch.equal(getattr(a, f).view(torch.uint8), getattr(b, f).view(torch.uint8)), f


def _code(fn):
    from voxelengine_amd import _lib
    with pytest.raises(_lib.VXError) as ei:
        fn()
    return ei.value.code


def test_context_allocation_failure(disarm):
    api = disarm
    from voxelengine_amd import _lib
    for k in (0, 1):                       # the context's two device buffers
        api.inject_failure(api.FAIL_ALLOC, k)
        assert _code(lambda: api.Context(0)) == _lib.VX_ERR_OUT_OF_MEMORY
    c = api.Context(0)
    c.close()


def test_volume_allocation_failure(ctx, disarm):
    api = disarm
    from voxelengine_amd import _lib
    for k in (0, 2, 3):                    # cells, a level, the march grid
        api.inject_failure(api.FAIL_ALLOC, k)
        assert _code(lambda: api.Volume(ctx, 32, 32, 32, nlevels=2)) == _lib.VX_ERR_OUT_OF_MEMORY
    v = api.Volume(ctx, 32, 32, 32, nlevels=2)
    v.clear()
    assert not v.download().any()
    v.close()


def test_trace_scratch_allocation_failure_then_retry(scene, disarm):
    """A fresh context grows its ray scratch on the first trace: the failed
    growth is reported and the retry traces the same bits."""
    api = disarm
    from voxelengine_amd import _lib
    s = scene
    c2 = api.Context(0)
    try:
        vis = api.Visibility(s["W"], s["H"], fetches=True)
        api.inject_failure(api.FAIL_ALLOC, 0)
        assert _code(lambda: api.trace(c2, s["occ"], s["vd"], s["g"], s["pal"], s["params"], vis)) == \
            _lib.VX_ERR_OUT_OF_MEMORY
        api.inject_failure(api.FAIL_ALLOC, 1)
        assert _code(lambda: api.trace(c2, s["occ"], s["vd"], s["g"], s["pal"], s["params"], vis)) == \
            _lib.VX_ERR_OUT_OF_MEMORY
        api.trace(c2, s["occ"], s["vd"], s["g"], s["pal"], s["params"], vis)
        torch.cuda.synchronize()
        _same(vis, s["ref"])
    finally:
        c2.close()


def test_trace_launch_failure_then_retry(ctx, scene, disarm):
    api = disarm
    from voxelengine_amd import _lib
    s = scene
    vis = api.Visibility(s["W"], s["H"], fetches=True)
    api.inject_failure(api.FAIL_LAUNCH, 0)
    assert _code(lambda: api.trace(ctx, s["occ"], s["vd"], s["g"], s["pal"], s["params"], vis)) == \
        _lib.VX_ERR_DEVICE
    torch.cuda.synchronize()
    api.trace(ctx, s["occ"], s["vd"], s["g"], s["pal"], s["params"], vis)
    torch.cuda.synchronize()
    _same(vis, s["ref"])


def test_rebuild_allocation_and_launch_failure_then_retry(scene, disarm):
    api = disarm
    from voxelengine_amd import _lib
    s = scene
    c2 = api.Context(0)
    try:
        d = s["model"].dims
        v = api.Volume(c2, d[0] // 2, d[1] // 2, d[2] // 2, nlevels=2)
        v.clear()
        inst = [api.make_instance(s["model"].mip_ptr(0), d, IDENTITY, IDENTITY)]
        api.inject_failure(api.FAIL_ALLOC, 0)
        assert _code(lambda: v.rebuild(inst)) == _lib.VX_ERR_OUT_OF_MEMORY
        api.inject_failure(api.FAIL_LAUNCH, 0)
        assert _code(lambda: v.rebuild(inst)) == _lib.VX_ERR_DEVICE
        torch.cuda.synchronize()
        v.clear()
        v.rebuild(inst)
        v.build_levels()
        torch.cuda.synchronize()
        assert np.array_equal(v.download(), s["occ"].download())
        v.close()
    finally:
        c2.close()


def _items(api, ctx, s, sizes):
    from voxelengine_amd import views
    items, plain = [], []
    for k, (W, H) in enumerate(sizes):
        vd = views.scene_view(s["model"].dims, W, H, variant=k, frame=2)
        g = api.GBuffer(W, H)
        api.primary(ctx, s["model"], vd, g)
        p = api.Visibility(W, H, fetches=True)
        api.trace(ctx, s["occ"], vd, g, s["pal"], s["params"], p)
        items.append((s["occ"], vd, g, api.Visibility(W, H, fetches=True)))
        plain.append(p)
    return items, plain


def test_batch_recaptures_on_shape_change(ctx, scene):
    from voxelengine_amd import api
    s = scene
    items, plain = _items(api, ctx, s, [(64, 48), (80, 40)])
    b = api.Batch(ctx, items, s["pal"], s["params"])
    try:
        assert b.info() == (2, 1)
        b.launch()
        torch.cuda.synchronize()
        for (_, _, _, o), p in zip(items, plain):
            _same(o, p)
        # more items, a larger view: new tile queue, larger scratch, new graph
        items2, plain2 = _items(api, ctx, s, [(32, 24), (200, 120), (72, 56)])
        b.update(items2)
        assert b.info() == (3, 2)
        for _ in range(2):
            b.launch()
            torch.cuda.synchronize()
            for (_, _, _, o), p in zip(items2, plain2):
                _same(o, p)
        # back to one small item
        items3, plain3 = _items(api, ctx, s, [(40, 40)])
        b.update(items3)
        assert b.info() == (1, 3)
        b.launch()
        b.launch_direct()
        torch.cuda.synchronize()
        _same(items3[0][3], plain3[0])
    finally:
        b.close()


def test_batch_failed_update_refuses_launch(ctx, scene, disarm):
    api = disarm
    from voxelengine_amd import _lib
    s = scene
    items, plain = _items(api, ctx, s, [(64, 48)])
    b = api.Batch(ctx, items, s["pal"], s["params"])
    try:
        # a bad item is refused before anything changes: the batch still launches
        bad = [(s["occ"], dict(items[0][1], width=0), items[0][2], items[0][3])]
        assert _code(lambda: b.update(bad)) == _lib.VX_ERR_INVALID_ARGUMENT
        assert b.info() == (1, 1)
        b.launch()
        torch.cuda.synchronize()
        _same(items[0][3], plain[0])
        # a launch failure while capturing leaves the batch unusable until re-planned
        items2, plain2 = _items(api, ctx, s, [(96, 64), (48, 32)])
        api.inject_failure(api.FAIL_LAUNCH, 0)
        assert _code(lambda: b.update(items2)) == _lib.VX_ERR_DEVICE
        assert _code(lambda: b.launch()) == _lib.VX_ERR_STATE
        assert _code(lambda: b.launch_direct()) == _lib.VX_ERR_STATE
        b.update(items2)
        assert b.info() == (2, 2)
        b.launch()
        torch.cuda.synchronize()
        for (_, _, _, o), p in zip(items2, plain2):
            _same(o, p)
    finally:
        b.close()


def test_environment_arms_the_allocation_failure():
    """VX_FAIL_ALLOC=1 in the environment: the second device allocation after
    the library loads (the context's queue counters) fails, the next context
    succeeds."""
    code = ("import torch\n"
            "from voxelengine_amd import api, _lib\n"
            "try:\n"
            "    api.Context(0)\n"
            "    print('no failure')\n"
            "except _lib.VXError as e:\n"
            "    print('code', e.code)\n"
            "c = api.Context(0)\n"
            "c.close()\n"
            "print('second context ok')\n")
    env = dict(os.environ, VX_FAIL_ALLOC="1")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "code -2" in r.stdout and "second context ok" in r.stdout, r.stdout
