"""Batched trace: many uneven scenes and views in one graph-captured launch must
equal the per-view launches bit for bit, on every graph replay, with eager
work and synchronizations between replays; one item is also checked against
the oracle."""
import numpy as np
import pytest
import torch

from helpers import IDENTITY, assert_bitwise, oracle_occ

pytestmark = pytest.mark.gpu


def _items(api, ctx, views, configs, n=6, seed=3):
    rng = np.random.default_rng(seed)
    items = []
    for k in range(n):
        dims = tuple(int(8 * rng.integers(4, 13)) for _ in range(3))
        model = api.Model(ctx, *dims)
        model.generate(seed=seed + k, terrain_fill=float(rng.uniform(0.1, 0.4)), nboxes=int(rng.integers(0, 30)),
                       box_min=2, box_max=10)
        model.build_mips()
        d = model.dims
        occ = api.Volume(ctx, d[0] // 2, d[1] // 2, d[2] // 2, nlevels=int(rng.integers(0, 3)))
        occ.clear()
        occ.rebuild([api.make_instance(model.mip_ptr(0), d, IDENTITY, IDENTITY)])
        occ.build_levels()
        w, h = int(rng.integers(9, 90)), int(rng.integers(9, 70))
        vd = views.scene_view(d, w, h, variant=k)
        g = api.GBuffer(w, h)
        api.primary(ctx, model, vd, g)
        items.append(dict(model=model, occ=occ, vd=vd, g=g, w=w, h=h))
    return items


def test_batch_equals_plain_launches_and_oracle(ctx, oracle):
    from voxelengine_amd import api, configs, views
    items = _items(api, ctx, views, configs)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_ao=6, ao_maxt=24.0, march="dda")
    plain = []
    for it in items:
        vis = api.Visibility(it["w"], it["h"], fetches=True)
        api.trace(ctx, it["occ"], it["vd"], it["g"], pal, params, vis)
        plain.append(vis)
    outs = [api.Visibility(it["w"], it["h"], fetches=True) for it in items]
    batch = api.Batch(ctx, [(it["occ"], it["vd"], it["g"], o) for it, o in zip(items, outs)], pal, params)
    for replay in range(4):
        for o in outs:
            o.ao.fill_(-1.0); o.spec.fill_(-1.0); o.shadow.fill_(7); o.fetches.fill_(-1)
        if replay == 2:
            batch.launch_direct()
        else:
            batch.launch()
        torch.cuda.synchronize()
        _ = torch.randn(1000, device="cuda").sum().item()      # eager work between replays
        for k, (o, p) in enumerate(zip(outs, plain)):
            for f in ("shadow", "ao", "spec", "fetches"):
                assert_bitwise(getattr(o, f).cpu().numpy(), getattr(p, f).cpu().numpy(), f"item {k} {f} replay {replay}")
    it = items[1]
    occ_h = oracle_occ(oracle, it["model"].download(0))
    pos, nrm, mat = it["g"].pos.cpu().numpy(), it["g"].normal.cpu().numpy(), it["g"].material.cpu().numpy()
    sh, ao, sp, fe = oracle.trace(occ_h, it["vd"], params, pos, nrm, mat, configs.make_palette())
    assert_bitwise(outs[1].ao.cpu().numpy(), ao, "oracle ao")
    assert_bitwise(outs[1].shadow.cpu().numpy(), sh, "oracle shadow")
    assert_bitwise(outs[1].spec.cpu().numpy(), sp, "oracle spec")
    batch.close()
    # the same batch without fetch counts: same bits
    outs2 = [api.Visibility(it["w"], it["h"]) for it in items]
    batch2 = api.Batch(ctx, [(it["occ"], it["vd"], it["g"], o) for it, o in zip(items, outs2)], pal, params)
    batch2.launch()
    torch.cuda.synchronize()
    for k, (o, p) in enumerate(zip(outs2, plain)):
        for f in ("shadow", "ao", "spec"):
            assert_bitwise(getattr(o, f).cpu().numpy(), getattr(p, f).cpu().numpy(), f"item {k} {f} without fetch counts")
    batch2.close()


def test_batch_rejects_reference_marchers(ctx):
    from voxelengine_amd import api, configs, _lib
    params = configs.trace_params(march="reference")
    pal = torch.from_numpy(configs.make_palette()).cuda()
    vol = api.Volume(ctx, 8, 8, 8, nlevels=1)
    g = api.GBuffer(8, 8)
    vis = api.Visibility(8, 8)
    from voxelengine_amd import views
    with pytest.raises(_lib.VXError):
        api.Batch(ctx, [(vol, views.scene_view((16, 16, 16), 8, 8), g, vis)], pal, params)


def test_every_autotuner_candidate_gives_the_same_bits(ctx, tmp_path):
    """Each tile shape, queue order and queue-items-per-tile setting the autotuner tries,
    for a batch and for the plain trace, equals the default launch bit for
    bit; tune_batch returns a working batch under the winning tuning."""
    from voxelengine_amd import api, autotune, configs, views
    items = _items(api, ctx, views, configs, n=5, seed=9)
    pal = torch.from_numpy(configs.make_palette()).cuda()
    params = configs.trace_params(n_ao=5, ao_maxt=24.0, march="dda")
    ref = []
    for it in items:
        vis = api.Visibility(it["w"], it["h"], fetches=True)
        api.trace(ctx, it["occ"], it["vd"], it["g"], pal, params, vis)
        ref.append(vis)
    try:
        for cand in autotune.CANDIDATES:
            ctx.set_tuning(**cand)
            outs = [api.Visibility(it["w"], it["h"], fetches=True) for it in items]
            for it, o in zip(items, outs):
                api.trace(ctx, it["occ"], it["vd"], it["g"], pal, params, o)
            if cand.get("kernel") == 2:
                bouts = [api.Visibility(it["w"], it["h"], fetches=True) for it in items]
                b = api.Batch(ctx, [(it["occ"], it["vd"], it["g"], o) for it, o in zip(items, bouts)], pal, params)
                b.launch()
                torch.cuda.synchronize()
                b.close()
                outs = outs + bouts
            torch.cuda.synchronize()
            for k, o in enumerate(outs):
                p = ref[k % len(items)]
                for f in ("shadow", "ao", "spec", "fetches"):
                    assert_bitwise(getattr(o, f).cpu().numpy(), getattr(p, f).cpu().numpy(), f"{cand} item {k} {f}")
        ctx.reset_tuning()
        outs = [api.Visibility(it["w"], it["h"]) for it in items]
        batch, best, times = autotune.tune_batch(
            ctx, lambda: api.Batch(ctx, [(it["occ"], it["vd"], it["g"], o) for it, o in zip(items, outs)], pal, params),
            "test-batch", cache_path=str(tmp_path / "tune.json"), reps=1)
        assert len(times) == len(autotune.BATCH_CANDIDATES) and best in autotune.BATCH_CANDIDATES
        batch.launch()
        torch.cuda.synchronize()
        for k, o in enumerate(outs):
            assert_bitwise(o.ao.cpu().numpy(), ref[k].ao.cpu().numpy(), f"tuned batch item {k}")
        batch.close()
    finally:
        ctx.reset_tuning()
