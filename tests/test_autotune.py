"""The autotuner's search and cache (voxelengine_amd/autotune.py), on the host
with a stand-in context and timer: the fastest candidate wins, the choice is
cached per workload key and reused without timing, and a batched search
closes every batch but the winner's."""
import json

from voxelengine_amd import autotune


class FakeCtx:
    def __init__(self):
        self.tuning = None

    def set_tuning(self, **kw):
        self.tuning = dict(kw)


def test_candidates_cover_tile_shapes_orders_and_chunks():
    c = autotune.MARCH_CANDIDATES
    assert {x["tile"] for x in c} == {0, 1, 2} and {x["order"] for x in c} == {0, 1}
    assert {x["chunks"] for x in c} == {0, 1, 3} and len(c) == 18
    # a batch takes one queue item per tile: its search skips the chunk knob
    assert autotune.BATCH_CANDIDATES == [x for x in c if x["chunks"] == 0] and len(autotune.BATCH_CANDIDATES) == 6
    assert autotune.CANDIDATES[0] == dict(kernel=1)


def test_tune_picks_the_fastest_and_caches(tmp_path):
    ctx = FakeCtx()
    cands = autotune.MARCH_CANDIDATES
    want = cands[11]
    timed = []

    def timer(run, reps):
        timed.append(dict(ctx.tuning))
        return 1.0 if ctx.tuning == want else 2.0 + 0.01 * len(timed)

    path = str(tmp_path / "tune.json")
    best, times = autotune.tune(ctx, lambda: None, "key-a", candidates=cands, cache_path=path, timer=timer)
    assert best == want and ctx.tuning == want and len(times) == len(cands) == len(timed)
    on_disk = json.load(open(path))
    assert on_disk["key-a"]["tuning"] == want and on_disk["key-a"]["candidates"] == cands

    def no_timer(run, reps):
        raise AssertionError("a cached key is not timed again")

    ctx2 = FakeCtx()
    best2, times2 = autotune.tune(ctx2, lambda: None, "key-a", candidates=cands, cache_path=path, timer=no_timer)
    assert best2 == want and ctx2.tuning == want and times2 == times
    # another key, or force, searches again
    best3, _ = autotune.tune(ctx2, lambda: None, "key-b", candidates=cands[:3], cache_path=path,
                             timer=lambda run, reps: 5.0 - reps * 0 - cands.index(ctx2.tuning))
    assert best3 == cands[2]
    assert set(json.load(open(path))) == {"key-a", "key-b"}


def test_tune_batch_returns_the_winners_batch(tmp_path):
    ctx = FakeCtx()
    made = []

    class Batch:
        def __init__(self):
            self.tuning = dict(ctx.tuning)
            self.closed = False
            made.append(self)

        def launch(self):
            pass

        def close(self):
            self.closed = True

    cands = autotune.MARCH_CANDIDATES[:5]
    b, best, times = autotune.tune_batch(ctx, Batch, "k", candidates=cands, cache_path=str(tmp_path / "t.json"),
                                         timer=lambda run, reps: 3.0 if run.__self__.tuning != cands[3] else 1.0)
    assert best == cands[3] and b.tuning == cands[3] and not b.closed
    assert all(m.closed for m in made[:-1]) and len(made) == len(cands) + 1


def test_workload_key_names_the_shape():
    from voxelengine_amd import configs
    p = configs.trace_params(n_ao=8, n_spec=0)
    k = autotune.workload_key((64, 72, 80), 2, 640, 360, p, device_name="dev")
    assert k == "dev|64x72x80|L2|640x360|1+8+0|m000"
