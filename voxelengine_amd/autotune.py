"""Autotuner for the ray-parallel trace's launch shape.

Times candidate vx_trace_tuning settings on the caller's own workload with
HIP events, keeps the fastest and caches the choice per workload key in a
JSON file (VX_TUNE_CACHE, default ~/.cache/voxelengine_amd/tune.json). The
knobs are the tile shape (8x8, 16x4 or 4x16 pixels per 64-ray batch), the
tile queue order (rows of tiles, or rows of 8x8-tile blocks so that the tiles
an XCD takes in turn cover a square screen region) and the queue items per
tile (the library's own choice from the launch's size, one item per tile, or
three); for a plain trace the per-pixel kernel is a candidate too. Every candidate computes identical bits; only the
time differs. A batched launch is captured with the tuning current when it
is created, so tune_batch creates one batch per candidate and returns the
fastest one.
"""
from __future__ import annotations

import itertools
import json
import os

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "voxelengine_amd", "tune.json")

# tile shape x queue order x queue items per tile (0: the library's choice)
MARCH_CANDIDATES = [dict(kernel=2, tile=t, order=o, chunks=c)
                    for t, o, c in itertools.product((0, 1, 2), (0, 1), (0, 1, 3))]
CANDIDATES = [dict(kernel=1)] + MARCH_CANDIDATES
BATCH_CANDIDATES = [c for c in MARCH_CANDIDATES if c["chunks"] == 0]


def workload_key(volume_dims, nlevels, width, height, params, device_name=None) -> str:
    if device_name is None:
        import torch
        device_name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    rays = f"{params['n_shadow']}+{params['n_ao']}+{params['n_spec']}"
    marchers = f"{params['shadow_march']}{params['ao_march']}{params['spec_march']}"
    return f"{device_name}|{'x'.join(map(str, volume_dims))}|L{nlevels}|{width}x{height}|{rays}|m{marchers}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, cache):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_launch(run, reps=3, warmup=1) -> float:
    """Average ms of run() (which enqueues one launch on the current stream)."""
    import torch
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _search(ctx, candidates, measure):
    times, best, best_ms = {}, None, float("inf")
    for i, cand in enumerate(candidates):
        ctx.set_tuning(**cand)
        ms = measure(cand)
        times[str(i)] = round(ms, 4)
        if ms < best_ms:
            best, best_ms = cand, ms
    return best, times


def _cached(cache_path, key, force):
    cache = _load(cache_path)
    if key in cache and not force:
        return cache, cache[key]
    return cache, None


def _store(cache_path, cache, key, best, times, candidates):
    cache[key] = {"tuning": best, "ms": times, "candidates": candidates}
    try:
        _save(cache_path, cache)
    except OSError:
        pass


def tune(ctx, run, key: str, candidates=None, cache_path=None, reps=3, force=False, timer=time_launch):
    """Sets ctx's tuning to the fastest candidate for `run` (one trace launch)
    and returns (tuning dict, {candidate index: ms})."""
    cache_path = cache_path or os.environ.get("VX_TUNE_CACHE", DEFAULT_CACHE)
    cache, hit = _cached(cache_path, key, force)
    if hit:
        ctx.set_tuning(**hit["tuning"])
        return hit["tuning"], hit.get("ms", {})
    candidates = candidates or CANDIDATES
    best, times = _search(ctx, candidates, lambda cand: timer(run, reps=reps))
    ctx.set_tuning(**best)
    _store(cache_path, cache, key, best, times, candidates)
    return best, times


def tune_batch(ctx, make_batch, key: str, candidates=None, cache_path=None, reps=3, force=False,
               timer=time_launch):
    """make_batch() creates a batch (an object with launch() and close()) under
    ctx's current tuning. Returns (batch, tuning dict, {candidate index: ms})
    for the fastest candidate; the others are closed."""
    cache_path = cache_path or os.environ.get("VX_TUNE_CACHE", DEFAULT_CACHE)
    cache, hit = _cached(cache_path, key, force)
    if hit:
        ctx.set_tuning(**hit["tuning"])
        return make_batch(), hit["tuning"], hit.get("ms", {})
    # a batch takes one queue item per tile whatever the chunk setting
    candidates = candidates or BATCH_CANDIDATES

    def measure(cand):
        b = make_batch()
        try:
            return timer(b.launch, reps=reps)
        finally:
            b.close()

    best, times = _search(ctx, candidates, measure)
    ctx.set_tuning(**best)
    _store(cache_path, cache, key, best, times, candidates)
    return make_batch(), best, times
