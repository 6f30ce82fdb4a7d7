"""Autotuner for the trace launch shape.

Measures candidate vx_trace_tuning settings (kernel variant, block size,
persistent blocks per CU, lane-refill threshold, DDA steps between refill
checks) on the caller's own workload with HIP events, keeps the fastest, and
caches the choice per workload key in a JSON file (VX_TUNE_CACHE, default
~/.cache/voxelengine_amd/tune.json). Every candidate computes identical bits;
only the time differs.
"""
from __future__ import annotations

import json
import os

import torch

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "voxelengine_amd", "tune.json")

CANDIDATES = [
    dict(kernel=1),
    dict(kernel=2, block=256, refill=16, steps=4),
    dict(kernel=2, block=256, refill=8, steps=4),
    dict(kernel=2, block=256, refill=32, steps=4),
    dict(kernel=2, block=256, refill=16, steps=2),
    dict(kernel=2, block=256, refill=16, steps=8),
    dict(kernel=2, block=512, refill=16, steps=4),
]


def workload_key(volume_dims, nlevels, width, height, params) -> str:
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    rays = f"{params['n_shadow']}+{params['n_ao']}+{params['n_spec']}"
    marchers = f"{params['shadow_march']}{params['ao_march']}{params['spec_march']}"
    return f"{dev}|{'x'.join(map(str, volume_dims))}|L{nlevels}|{width}x{height}|{rays}|m{marchers}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, cache):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_launch(run, reps=3, warmup=1) -> float:
    """Average ms of run() (which enqueues one launch on the current stream)."""
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


def tune(ctx, run, key: str, candidates=None, cache_path=None, reps=3, force=False):
    """Sets ctx's tuning to the fastest candidate for `run` and returns
    (tuning dict, {candidate index: ms})."""
    cache_path = cache_path or os.environ.get("VX_TUNE_CACHE", DEFAULT_CACHE)
    cache = _load(cache_path)
    if key in cache and not force:
        ctx.set_tuning(**cache[key]["tuning"])
        return cache[key]["tuning"], cache[key].get("ms", {})
    candidates = candidates or CANDIDATES
    times = {}
    best, best_ms = None, float("inf")
    for i, cand in enumerate(candidates):
        ctx.set_tuning(**cand)
        ms = time_launch(run, reps=reps)
        times[str(i)] = round(ms, 4)
        if ms < best_ms:
            best, best_ms = cand, ms
    ctx.set_tuning(**best)
    cache[key] = {"tuning": best, "ms": times}
    try:
        _save(cache_path, cache)
    except OSError:
        pass
    return best, times
