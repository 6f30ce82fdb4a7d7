"""vx_trace_noalloc (the trace with no allocation) against vx_trace through the
C ABI: identical visibility and fetch counts on two calls in a row, the second
on a smaller view, so the shared scratch keeps the first call's contents."""
import ctypes as C

import numpy as np
import pytest

from helpers import device_scene

pytestmark = pytest.mark.gpu


def _call(L, name, ctx, occ, vd, g, pal, tp, vis, stream):
    from voxelengine_amd import api
    v = api.view_struct(vd)
    gs, o = g.struct(), vis.struct()
    return getattr(L, name)(ctx.h, occ.h, C.byref(v), C.byref(gs), api._ptr(pal), C.byref(tp), C.byref(o), stream)


@pytest.mark.parametrize("ao", [16, 64])
def test_noalloc_trace_equals_trace(ctx, ao):
    import torch
    from voxelengine_amd import _lib, api, configs, views
    L = _lib.lib()
    model, occ, vol = device_scene(api, ctx, 96, seed=21, nboxes=40)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_ao=ao, ao_maxt=32.0, n_spec=1, march="dda")
    tp = api.trace_params(params)
    stream = torch.cuda.current_stream().cuda_stream
    # the second call is smaller; at 64 AO rays the first spans several tile ranges
    for (W, H), fetches in (((400, 300), True), ((123, 77), False)):
        vd = views.scene_view((96, 96, 96), W, H)
        g = api.GBuffer(W, H)
        api.primary(ctx, model, vd, g)
        a = api.Visibility(W, H, fetches=fetches)
        b = api.Visibility(W, H, fetches=fetches)
        assert _call(L, "vx_trace", ctx, occ, vd, g, pal, tp, a, stream) == 0
        assert _call(L, "vx_trace_noalloc", ctx, occ, vd, g, pal, tp, b, stream) == 0, L.vx_last_error()
        torch.cuda.synchronize()
        hit = int((g.material > 0).sum().item())
        assert hit > 0.3 * W * H
        assert torch.equal(a.shadow, b.shadow)
        assert torch.equal(a.ao.view(torch.int32), b.ao.view(torch.int32))
        assert torch.equal(a.spec.view(torch.int32), b.spec.view(torch.int32))
        if fetches:
            assert torch.equal(a.fetches, b.fetches) and int(a.fetches.sum().item()) > 0


def test_noalloc_trace_replays_in_a_graph(ctx):
    """The caller it is for: a HIP graph. vx_trace_noalloc captured on a side
    stream (torch.cuda.graph) and replayed three times, with eager work and
    synchronizations between the replays and the outputs overwritten before
    each, gives vx_trace's bits every time."""
    import torch
    from voxelengine_amd import _lib, api, configs, views
    L = _lib.lib()
    model, occ, vol = device_scene(api, ctx, 96, seed=22, nboxes=40)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_ao=8, ao_maxt=32.0, n_spec=1, march="dda")
    tp = api.trace_params(params)
    W, H = 160, 96
    vd = views.scene_view((96, 96, 96), W, H)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, vd, g)
    ref = api.Visibility(W, H)
    main = torch.cuda.current_stream().cuda_stream
    assert _call(L, "vx_trace", ctx, occ, vd, g, pal, tp, ref, main) == 0
    out = api.Visibility(W, H)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        # one eager call on the side stream first (the launch shape is then cached)
        assert _call(L, "vx_trace_noalloc", ctx, occ, vd, g, pal, tp, out, side.cuda_stream) == 0
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        assert _call(L, "vx_trace_noalloc", ctx, occ, vd, g, pal, tp, out, side.cuda_stream) == 0, L.vx_last_error()
    torch.cuda.synchronize()
    for k in range(3):
        out.shadow.fill_(7)
        out.ao.fill_(-1.0)
        out.spec.fill_(-1.0)
        _ = torch.randn(4096, device="cuda").sum().item()    # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.shadow, ref.shadow), k
        assert torch.equal(out.ao.view(torch.int32), ref.ao.view(torch.int32)), k
        assert torch.equal(out.spec.view(torch.int32), ref.spec.view(torch.int32)), k
    hit = int((g.material > 0).sum().item())
    assert hit > 0.3 * W * H
