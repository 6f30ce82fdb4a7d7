"""Benchmark: secondary-ray throughput of the voxel-traversal lighting path on one MI355X.

Workload (default): 512^3 terrain + 2000 boxes, 2 coarse occupancy levels, one
1920x1080 view; per hit pixel 1 shadow ray (maxT 128), 16 AO rays (length 32)
and 1 GGX specular-occlusion ray (maxT 256), all through the DDA marcher
(raycastShadowVolume semantics, bit-exact). A step is one vx_trace launch over
the view; inputs (volume, levels, G-buffer) are resident in HBM before timing.

Prints one JSON line. Extra fields: the reference-marcher variant of the same
frame, the backward (shading adjoint) rate, rebuild and primary-hit times,
the point/spot light pass (64 + 64 lights), the reflection pass and the
screen-space occlusion over the same frame, config B
(batched against plain launches), the prefabs (load, rebuild, one frame with
geometry pass, trace and temporal accumulation), config L, the HBM roofline
of the trace kernel and the CPU oracle baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PIXEL_IO_BYTES = 34   # hit pos 12 + normal 12 + material 1 in; shadow 1 + AO 4 + spec 4 out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


KERNEL_SRC = os.path.join(ROOT, "voxelengine_amd", "csrc", "vx_kernels.hip")
# the bench re-tunes every run (force=True); the cache only records the choice
TUNE_CACHE = os.path.join(tempfile.gettempdir(), "vx_bench_tune.json")
COUNTERS = os.path.join(ROOT, "profiles", "k_march_F_1080p_512_counters.json")
MI355X_SIMDS = 1024          # 256 CUs x 4 SIMDs
MI355X_CLOCK_HZ = 2.4e9      # peak engine clock
# Issue ceilings in wave instructions per SIMD per cycle, measured on the
# device by voxelengine_amd/valu_peak (independent chains of v_add_f32 /
# v_add_u32 / s_add_u32, 8 waves per SIMD; rocprofv3 SQ_INSTS_* over
# GRBM_GUI_ACTIVE cycles): a wave64 vector instruction every 2.25 cycles, a
# scalar one every 4.1.
VALU_ISSUE_PER_CYCLE = 0.444
SALU_ISSUE_PER_CYCLE = 0.244


def kernel_src_sha16():
    import hashlib
    with open(KERNEL_SRC, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def profiled_counters():
    """rocprofv3 counters of k_march on this workload (HBM traffic from the
    FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel of known size, L2
    hits and misses, vector / scalar instructions), summarized by
    `bench.py --summarize-profiles` into the committed counters file. Returned
    only when they were measured on the kernel source of this tree (its
    sha256 prefix is stored with them); otherwise None."""
    try:
        with open(COUNTERS) as f:
            c = json.load(f)
    except (OSError, ValueError):
        return None
    return c if c.get("kernel_src_sha16") == kernel_src_sha16() else None


def cpu_baseline(occ_host, vd, params, pos, nrm, mat, pal, budget_s=12.0):
    """Oracle (C restatement, OpenMP over pixels) on a bounded, deterministic
    sample of the same frame: every k-th hit pixel, sized from a pilot run to
    about budget_s; when the whole frame takes less, the whole frame repeated."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    hit = np.nonzero(mat > 0)[0].astype(np.int32)
    pilot = hit[:: max(1, hit.size // 2048)]
    t0 = time.perf_counter()
    O.trace(occ_host, vd, params, pos, nrm, mat, pal, pixels=pilot, nthreads=cores)
    dt = time.perf_counter() - t0
    n_target = int(min(hit.size, max(pilot.size, pilot.size * budget_s / max(dt, 1e-6))))
    stride = max(1, hit.size // n_target)
    sample = np.ascontiguousarray(hit[::stride])
    reps, t_total = 0, 0.0
    while reps < 64 and (reps == 0 or t_total < budget_s):
        t0 = time.perf_counter()
        O.trace(occ_host, vd, params, pos, nrm, mat, pal, pixels=sample, nthreads=cores)
        t_total += time.perf_counter() - t0
        reps += 1
    rays = reps * sample.size * (params["n_shadow"] + params["n_ao"] + params["n_spec"])
    what = "the whole frame" if stride == 1 else f"every {stride}th hit pixel"
    return {"value": rays / t_total / 1e9, "unit": "Grays/s", "cores": cores, "kind": "port",
            "label": "C++ restatement (not lavapipe): the oracle, C with OpenMP over pixels",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{what} ({sample.size} px) x {reps}: {rays} rays in {t_total:.1f} s"}


def limiter(counters, kernel_ms):
    """What the rocprofv3 counters of this kernel source show bounds k_march,
    each as a fraction of the counted launch's time (1.0 = that resource busy
    for the whole launch): vector and scalar instruction issue (SQ_INSTS_VALU /
    SQ_INSTS_SALU wave instructions over 1024 SIMDs at the measured engine
    clock, against the issue ceilings measured on the device), the
    vector-memory address and data paths (TA / TD busy cycles per CU; busy
    includes waiting for data), waves waiting on memory counters
    (SQ_WAIT_ANY / SQ_WAVE_CYCLES) and measured HBM traffic against the HBM
    roofline. `bound` names the largest. None without counters."""
    if not counters:
        return None
    # the fractions of the counted dispatch itself (its duration and clock),
    # so that they stay consistent when this run's box clocks differently
    t = (counters.get("kernel_ns_in_ta_pass") or kernel_ms * 1e6) * 1e-9
    clock = counters.get("engine_clock_hz") or MI355X_CLOCK_HZ
    cyc = t * clock
    fr = {"valu_issue": counters["sq_insts_valu_per_launch"] / VALU_ISSUE_PER_CYCLE / MI355X_SIMDS / cyc,
          "salu_issue": counters["sq_insts_salu_per_launch"] / SALU_ISSUE_PER_CYCLE / MI355X_SIMDS / cyc,
          "hbm": counters["traffic_bytes_per_launch"] / t / (HBM_PEAK_GBS * 1e9)}
    if "ta_busy_per_cu" in counters:
        fr["ta_busy"] = counters["ta_busy_per_cu"] / cyc
        fr["td_busy"] = counters["td_busy_per_cu"] / cyc
    out = {"bound": max(fr, key=fr.get)}
    out.update({k + "_frac": round(v, 4) for k, v in fr.items()})
    out.update({"engine_clock_ghz": round(clock / 1e9, 3), "l2_hit_rate": counters["l2_hit_rate"],
                "l1_hit_rate": counters.get("l1_hit_rate"), "mem_wait_frac": counters.get("wait_frac_mem"),
                "issue_wait_frac": counters.get("wait_frac"),
                "counters": os.path.relpath(COUNTERS, ROOT) + " @ kernel source " + counters["kernel_src_sha16"]})
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def bench_batched(ctx, pal_d, reps=5):
    """64 scenes of uneven size (each axis a multiple of 8 in [64,512]), one view
    each (256x256..1920x1080), 1 shadow + 8 AO rays: one graph-captured
    persistent launch against 64 plain per-view launches."""
    import torch
    from voxelengine_amd import api, autotune, configs, views
    params = configs.trace_params(n_shadow=1, n_ao=8, ao_maxt=32.0, n_spec=0, march="dda")
    items = []
    for sc in configs.batched_scenes():
        m = api.Model(ctx, *sc["dims"])
        m.generate(seed=sc["seed"], terrain_fill=sc["fill"], nboxes=sc["nboxes"], box_min=sc["box"][0],
                   box_max=sc["box"][1])
        m.build_mips()
        d = m.dims
        occ = api.Volume(ctx, d[0] // 2, d[1] // 2, d[2] // 2, nlevels=2)
        occ.clear()
        occ.rebuild([api.make_instance(m.mip_ptr(0), d, api.IDENTITY, api.IDENTITY)])
        occ.build_levels()
        vd = views.scene_view(d, sc["width"], sc["height"], variant=sc["variant"])
        g = api.GBuffer(sc["width"], sc["height"])
        api.primary(ctx, m, vd, g)
        items.append((m, occ, vd, g, api.Visibility(sc["width"], sc["height"])))
    torch.cuda.synchronize()
    hit = sum(int((it[3].material > 0).sum().item()) for it in items)
    rays = hit * configs.rays_per_pixel(params)
    bvis = [api.Visibility(it[2]["width"], it[2]["height"]) for it in items]

    def plain():
        for (m, occ, vd, g, vis) in items:
            api.trace(ctx, occ, vd, g, pal_d, params, vis)

    # the autotuner picks the launch shape of each side on this workload:
    # one tuning for the 64 plain launches, one for the batched graph
    cache = TUNE_CACHE
    tuned_plain, plain_ms = autotune.tune(ctx, plain, "bench-B-plain", cache_path=cache, reps=2, force=True)
    batch, tuned_batch, batch_ms = autotune.tune_batch(
        ctx, lambda: api.Batch(ctx, [(it[1], it[2], it[3], bv) for it, bv in zip(items, bvis)], pal_d, params),
        "bench-B-batched", cache_path=cache, reps=2, force=True)

    res = {}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, fn, tu in (("plain_64_launches", plain, tuned_plain), ("batched_graph", batch.launch, tuned_batch)):
        ctx.set_tuning(**tu)
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 3), "grays_per_s": round(rays / (ms * 1e-3) / 1e9, 4), "tuning": tu}
    same = all(torch.equal(a[4].ao.view(torch.int32), b.ao.view(torch.int32)) and torch.equal(a[4].shadow, b.shadow)
               for a, b in zip(items, bvis))
    res["rays"] = rays
    res["batched_vs_plain"] = round(res["plain_64_launches"]["ms"] / res["batched_graph"]["ms"], 3)
    res["identical_outputs"] = bool(same)
    res["autotune_ms"] = {"plain": plain_ms, "batched": batch_ms}
    batch.close()
    ctx.reset_tuning()
    return res


def bench_large(ctx, pal_d, reps=3):
    """Config L: 1024^3 city blocks (20000 grounded boxes of edges 8..56, 5.2 %
    fill), 3 levels, four 3840x2160 views from inside the city with 1 shadow +
    32 AO + 1 spec ray per hit pixel in one graph-captured launch; palette and
    light gradients of all four views accumulated in one adjoint launch."""
    import torch
    from voxelengine_amd import api, autotune, configs, views
    cfg = configs.CONFIGS["L"]
    nx, ny, nz = cfg["dims"]
    W, H = cfg["width"], cfg["height"]
    params = configs.trace_params(**cfg["params"])
    m = api.Model(ctx, nx, ny, nz)
    m.generate(seed=configs.SEED, terrain_fill=cfg["fill"], nboxes=cfg["nboxes"], box_min=cfg["box"][0],
               box_max=cfg["box"][1], grounded=cfg["grounded"])
    m.build_mips()
    occ = api.Volume(ctx, nx // 2, ny // 2, nz // 2, nlevels=cfg["levels"])
    occ.clear()
    occ.rebuild([api.make_instance(m.mip_ptr(0), m.dims, api.IDENTITY, api.IDENTITY)])
    occ.build_levels()
    items = []
    for k in range(cfg["nviews"]):
        vd = configs.config_view(cfg, variant=k)
        g = api.GBuffer(W, H)
        api.primary(ctx, m, vd, g)
        items.append((occ, vd, g, api.Visibility(W, H)))
    torch.cuda.synchronize()
    hit = sum(int((it[2].material > 0).sum().item()) for it in items)
    rays = hit * configs.rays_per_pixel(params)
    batch, tuned, tune_ms = autotune.tune_batch(
        ctx, lambda: api.Batch(ctx, items, pal_d, params), "bench-L", reps=1, force=True,
        cache_path=TUNE_CACHE,
        candidates=autotune.BATCH_CANDIDATES)
    batch.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        batch.launch()
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    light = configs.default_light()
    gen = torch.Generator(device="cuda").manual_seed(configs.SEED)
    sitems = [(vd, g, vis, torch.randn((W * H, 3), generator=gen, device="cuda"), None) for (_, vd, g, vis) in items]
    dpal = torch.zeros((256, 6), dtype=torch.float32, device="cuda")
    dlight = torch.zeros(6, dtype=torch.float32, device="cuda")
    api.shade_backward_batch(ctx, sitems, pal_d, light, dpal, dlight)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        api.shade_backward_batch(ctx, sitems, pal_d, light, dpal, dlight)
    e1.record()
    torch.cuda.synchronize()
    bwd_ms = e0.elapsed_time(e1) / 5
    batch.close()
    ctx.reset_tuning()
    out = {"views": cfg["nviews"], "resolution": f"{W}x{H}", "hit_pixels": hit, "tuning": tuned,
           "hit_fraction": round(hit / (cfg["nviews"] * W * H), 4), "rays": rays,
           "forward_ms": round(fwd_ms, 3), "forward_grays_per_s": round(rays / (fwd_ms * 1e-3) / 1e9, 3),
           "backward_ms": round(bwd_ms, 4), "backward_grays_per_s": round(rays / (bwd_ms * 1e-3) / 1e9, 2),
           "backward_px_per_s": round(hit / (bwd_ms * 1e-3), 1)}
    del items, sitems, occ, m
    torch.cuda.empty_cache()
    return out


def _prefab_frame(ctx, P, models, vol, reps=10):
    """One 1920x1080 frame of a loaded prefab, the frame after its spawn (every
    entity's previous world matrix is its world matrix): the multi-instance
    geometry pass over its draws, 1 shadow + 16 AO + 1 spec DDA rays per hit
    pixel over that G-buffer against the prefab's rebuilt volume, the shading
    and the temporal accumulation."""
    import torch
    from voxelengine_amd import api, configs, views
    W, H = 1920, 1080
    vd = views.make_view((8.0, 16.0, 6.0), (32.0, 0.0, 32.0), W, H)
    cmds = P.vox_cmds(models, prev=[P.world(i) for i in range(len(P.entities))])
    g = api.GBuffer(W, H)
    ext = api.GBufferExt(W, H)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    api.primary_instances(ctx, cmds, vd, g, ext)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        api.primary_instances(ctx, cmds, vd, g, ext)
    e1.record()
    torch.cuda.synchronize()
    geo_ms = e0.elapsed_time(e1) / reps
    vol.build_levels()
    hit = int((g.material > 0).sum().item())
    params = configs.trace_params(**configs.CONFIGS["F"]["params"])
    rays = hit * configs.rays_per_pixel(params)
    pal_d = torch.from_numpy(configs.make_palette()).cuda()
    vis = api.Visibility(W, H)
    api.trace(ctx, vol, vd, g, pal_d, params, vis)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        api.trace(ctx, vol, vd, g, pal_d, params, vis)
    e1.record()
    torch.cuda.synchronize()
    tr_ms = e0.elapsed_time(e1) / reps
    # shading of the frame, then the temporal accumulation over it (history fed back)
    rgb = torch.zeros((W * H, 3), dtype=torch.float32, device="cuda")
    api.shade_forward(ctx, vd, g, vis, pal_d, configs.default_light(), rgb)
    hist = [torch.zeros((W * H, 4), dtype=torch.float32, device="cuda") for _ in range(2)]
    for k in range(3):
        api.light_taa(ctx, vd, rgb, hist[k % 2], hist[(k + 1) % 2], ext.depth, g.normal, g.material, pal_d,
                      motion=ext.motion, seed=1)
    torch.cuda.synchronize()
    e0.record()
    for k in range(reps):
        api.light_taa(ctx, vd, rgb, hist[k % 2], hist[(k + 1) % 2], ext.depth, g.normal, g.material, pal_d,
                      motion=ext.motion, seed=1)
    e1.record()
    torch.cuda.synchronize()
    taa_ms = e0.elapsed_time(e1) / reps
    return {"draws": len(cmds), "hit_pixels": hit, "geometry_ms": round(geo_ms, 3),
            "geometry_mpx_per_s": round(W * H / (geo_ms * 1e-3) / 1e6, 1), "trace_ms": round(tr_ms, 3),
            "trace_grays_per_s": round(rays / (tr_ms * 1e-3) / 1e9, 3), "taa_ms": round(taa_ms, 3),
            "taa_mpx_per_s": round(W * H / (taa_ms * 1e-3) / 1e6, 1)}


def bench_lights(ctx, occ, vd, g, pal_d, dims, reps=10):
    """The point/spot light pass over the headline frame with 64 point and 64
    spot lights, the reference's limit per type (LightPointPipeline.h:15,
    LightSpotPipeline.h), marched with the reference's shipped marchers
    (Sparse for point, SuperSparse for spot). A shadow ray is marched for every
    (hit pixel, light in range) pair; that count comes from the same test on
    the host in float64, and the fetch count from one counting launch."""
    import torch
    from voxelengine_amd import api, configs
    rng = np.random.default_rng(configs.SEED + 7)
    ext = np.array(dims, np.float64) * 0.1                   # world extent of the scene
    L = np.zeros(128, api.LIGHT_DTYPE)
    for k in range(L.size):
        L[k]["type"] = 0 if k < 64 else 1
        L[k]["position"] = rng.uniform([0.0, 0.3, 0.0], [1.0, 0.7, 1.0]) * ext
        L[k]["range"] = rng.uniform(0.1, 0.3) * ext.max()
        L[k]["color"] = rng.uniform(0.2, 4.0, 3)
        L[k]["attenuation"] = rng.uniform(0.5, 2.0)
        d = rng.normal(size=3)
        d[1] = -abs(d[1])
        L[k]["direction"] = d / np.linalg.norm(d)
        L[k]["angle"] = rng.uniform(0.2, 0.9)
        L[k]["angle_attenuation"] = rng.uniform(0.5, 3.0)
    lights = api.lights_to_device(L)
    npx = vd["width"] * vd["height"]
    hit = (g.material > 0)
    wpos = g.pos[hit].double() * 0.1
    lp = torch.from_numpy(np.ascontiguousarray(L["position"], np.float64)).cuda()
    lr = torch.from_numpy(np.ascontiguousarray(L["range"], np.float64)).cuda()
    rays = 0
    for k0 in range(0, L.size, 16):
        dist = torch.cdist(wpos, lp[k0:k0 + 16])
        rays += int((dist <= lr[k0:k0 + 16]).sum().item())
    rgb = torch.zeros((npx, 3), dtype=torch.float32, device="cuda")
    fetches = torch.zeros(npx, dtype=torch.int32, device="cuda")
    api.light_pass(ctx, occ, vd, g, pal_d, lights, L.size, rgb, seed=3, frame=1, fetches=fetches)
    torch.cuda.synchronize()
    s_ref = int(fetches.to(torch.int64).sum().item())
    # algorithmic bytes: one byte per reference fetch, plus per hit pixel the
    # G-buffer read (25 B) and the rgb read and write (24 B)
    alg = s_ref + 49 * int(hit.sum().item())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = {}
    for name, path in (("pixel_loop", 0), ("ray_list", 1)):
        ctx.set_tuning(lights=path)
        api.light_pass(ctx, occ, vd, g, pal_d, lights, L.size, rgb, seed=3, frame=1)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            api.light_pass(ctx, occ, vd, g, pal_d, lights, L.size, rgb, seed=3, frame=1)
        e1.record()
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1) / reps
    ctx.reset_tuning()
    ms = times["pixel_loop"]
    return {"point_lights": 64, "spot_lights": 64, "marchers": "sparse (point) / supersparse (spot)",
            "kernel": "pixel loop (one thread per pixel over its tile's culled lights; the ray list beside it)",
            "ms": round(ms, 4), "shadow_rays": rays, "grays_per_s": round(rays / (ms * 1e-3) / 1e9, 3),
            "ref_fetches": s_ref, "achieved_gbps": round(alg / (ms * 1e-3) / 1e9, 1),
            "frac_of_hbm": round(alg / (ms * 1e-3) / 8.0e12, 4),
            "ray_list_ms": round(times["ray_list"], 4),
            "ray_list_grays_per_s": round(rays / (times["ray_list"] * 1e-3) / 1e9, 3)}


def bench_reflection_ssao(ctx, occ, vd, g, pal_d, reps=10):
    """The reflection pass (LightReflection.frag: the reference's Sparse
    marcher to 256, reprojection into a synthetic light buffer, procedural
    sky) and the screen-space occlusion (calculateOcclusion) over the headline
    frame, from the G-buffer's D24 depth attachment. One reflection ray is
    marched per hit pixel whose depth is below 0.999."""
    import torch
    from voxelengine_amd import api
    npx = vd["width"] * vd["height"]
    depth = torch.zeros(npx, dtype=torch.float32, device="cuda")
    api.gbuf_depth(ctx, vd, g, depth, volume_rid=1)
    gen = torch.Generator(device="cuda").manual_seed(11)
    light = torch.rand((npx, 4), dtype=torch.float32, device="cuda", generator=gen)
    out = torch.zeros((npx, 4), dtype=torch.float32, device="cuda")
    t = torch.zeros(npx, dtype=torch.float32, device="cuda")
    ss = torch.zeros(npx, dtype=torch.float32, device="cuda")
    api.reflection(ctx, occ, vd, g, pal_d, light, depth, out, t=t, seed=3, frame=1)
    torch.cuda.synchronize()
    rays = int((~torch.isnan(t)).sum().item())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        api.reflection(ctx, occ, vd, g, pal_d, light, depth, out, seed=3, frame=1)
    e1.record()
    torch.cuda.synchronize()
    refl_ms = e0.elapsed_time(e1) / reps
    api.ssao(ctx, vd, g, depth, ss, seed=3)
    e0.record()
    for _ in range(reps):
        api.ssao(ctx, vd, g, depth, ss, seed=3)
    e1.record()
    torch.cuda.synchronize()
    ssao_ms = e0.elapsed_time(e1) / reps
    return {"reflection": {"march": "sparse", "maxt": 256.0, "sky": "procedural", "rays": rays,
                           "ms": round(refl_ms, 4), "grays_per_s": round(rays / (refl_ms * 1e-3) / 1e9, 3)},
            "ssao": {"samples_per_pixel": 4, "ms": round(ssao_ms, 4),
                     "mpx_per_s": round(npx / (ssao_ms * 1e-3) / 1e6, 1),
                     "occluded_fraction": round(float((ss < 1).float().mean().item()), 4)}}


def bench_prefabs(ctx, reps=5):
    """The reference's FarmHouse and ModernHouse prefabs (files rebuilt from
    tests/golden/prefabs.npz) through the loaders into the 524x188x524 volume:
    spawn (host), model loads (read + upload + mips) and the first-frame
    rebuild in the reference's order, timed on the host around each call."""
    import tempfile
    import torch
    from voxelengine_amd import api
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from golden_mods import write_mod_folder
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        mods = write_mod_folder(tmp)
        for name in ("FarmHouse", "ModernHouse"):
            t0 = time.perf_counter()
            for _ in range(reps):
                P = api.Prefab(mods, f"default/{name}.pf")
            spawn_ms = (time.perf_counter() - t0) / reps * 1e3
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            models = P.load_models(ctx)
            torch.cuda.synchronize()
            load_ms = (time.perf_counter() - t0) * 1e3
            insts = P.instances(models)
            vol = api.Volume(ctx, 524, 188, 524, nlevels=2)
            times = []
            for _ in range(reps):
                vol.clear()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                vol.rebuild(insts)
                torch.cuda.synchronize()
                times.append((time.perf_counter() - t0) * 1e3)
            voxels = sum(int(np.prod(m.dims)) for m in (models[P.entities[i].model] for i in P.rebuild_order()
                                                        if P.entities[i].model >= 0))
            res[name] = {"entities": len(P.entities), "instances": len(insts), "model_voxels": voxels,
                         "spawn_ms": round(spawn_ms, 3), "model_load_ms": round(load_ms, 2),
                         "rebuild_ms": round(min(times), 3)}
            if name == "FarmHouse":
                res[name]["frame_1080p"] = _prefab_frame(ctx, P, models, vol)
            vol.close()
            for m in models.values():
                m.close()
            P.close()
    return res


def _counter_rows(path):
    """{dispatch id: (kernel name, {counter: value})} of a rocprofv3
    run_counter_collection.csv."""
    import csv
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = out.setdefault(int(r["Dispatch_Id"]), (r["Kernel_Name"], {}))
            d[1][r["Counter_Name"]] = d[1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def summarize_profiles(d):
    """Per-launch k_march counters from rocprofv3 passes over
    `bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-extras --copy-probe`
    under d: pmc_fetch (FETCH_SIZE), pmc_write (WRITE_SIZE), pmc_tcc
    (TCC_HIT_sum TCC_MISS_sum), pmc_sq (SQ_INSTS_VALU SQ_INSTS_SALU
    SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_SCA SQ_WAIT_INST_ANY SQ_WAIT_ANY
    SQ_WAVE_CYCLES SQ_BUSY_CYCLES) and pmc_ta (TA_TA_BUSY_sum TD_TD_BUSY_sum
    TCP_TOTAL_CACHE_ACCESSES_sum TCP_TCC_READ_REQ_sum GRBM_GUI_ACTIVE). The
    timed launches (k_march without fetch counts, the first of them excluded:
    cold caches) are averaged. FETCH_SIZE / WRITE_SIZE are scaled by what they
    read for the 1 GiB device copy of --copy-probe; the engine clock is
    GRBM_GUI_ACTIVE (summed over the 8 XCDs) over the launch's own duration
    in the same pass. Writes the counters file and returns it."""
    # the single-view march of the default launch shape (256 threads, 16
    # steps between refill checks), without fetch counts
    march = "k_march<256, 16, true, false>"

    def per_launch(sub, names):
        rows = _counter_rows(os.path.join(d, sub, "run_counter_collection.csv"))
        ks = [v for _, (k, v) in sorted(rows.items()) if march in k][1:]
        copies = [v for _, (k, v) in sorted(rows.items()) if "copyBuffer" in k]
        avg = {n: sum(v[n] for v in ks) / len(ks) for n in names}
        return avg, copies, len(ks)

    f, fc, n1 = per_launch("pmc_fetch", ["FETCH_SIZE"])
    w, wc, n2 = per_launch("pmc_write", ["WRITE_SIZE"])
    t, _, n3 = per_launch("pmc_tcc", ["TCC_HIT_sum", "TCC_MISS_sum"])
    q, _, n4 = per_launch("pmc_sq", ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_SCA",
                                     "SQ_WAIT_INST_ANY", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES"])
    a, _, n5 = per_launch("pmc_ta", ["TA_TA_BUSY_sum", "TD_TD_BUSY_sum", "TCP_TOTAL_CACHE_ACCESSES_sum",
                                     "TCP_TCC_READ_REQ_sum", "GRBM_GUI_ACTIVE"])
    dur = _dispatch_ns(os.path.join(d, "pmc_ta", "run_counter_collection.csv"), march)
    gib = float(1 << 30)
    fscale = gib / (max(c["FETCH_SIZE"] for c in fc) * 1024.0)
    wscale = gib / (max(c["WRITE_SIZE"] for c in wc) * 1024.0)
    fetch_b = f["FETCH_SIZE"] * 1024.0 * fscale
    write_b = w["WRITE_SIZE"] * 1024.0 * wscale
    clock = a["GRBM_GUI_ACTIVE"] / 8.0 / (dur * 1e-9) if dur else None
    c = {"kernel": march, "workload": "F: 512^3, 1920x1080, 1 shadow + 16 AO(32) + 1 spec DDA rays per hit pixel",
         "kernel_src_sha16": kernel_src_sha16(), "launches_averaged": [n1, n2, n3, n4, n5],
         "fetch_size_scale": round(fscale, 4), "write_size_scale": round(wscale, 4),
         "hbm_fetch_bytes_per_launch": int(fetch_b), "hbm_write_bytes_per_launch": int(write_b),
         "traffic_bytes_per_launch": int(fetch_b + write_b),
         "tcc_hit_per_launch": int(t["TCC_HIT_sum"]), "tcc_miss_per_launch": int(t["TCC_MISS_sum"]),
         "l2_hit_rate": round(t["TCC_HIT_sum"] / (t["TCC_HIT_sum"] + t["TCC_MISS_sum"]), 4),
         "l1_accesses_per_launch": int(a["TCP_TOTAL_CACHE_ACCESSES_sum"]),
         "l1_hit_rate": round(1.0 - a["TCP_TCC_READ_REQ_sum"] / a["TCP_TOTAL_CACHE_ACCESSES_sum"], 4),
         "sq_insts_valu_per_launch": int(q["SQ_INSTS_VALU"]), "sq_insts_salu_per_launch": int(q["SQ_INSTS_SALU"]),
         "wait_frac": round(q["SQ_WAIT_INST_ANY"] / q["SQ_WAVE_CYCLES"], 4),
         "wait_frac_mem": round(q["SQ_WAIT_ANY"] / q["SQ_WAVE_CYCLES"], 4),
         "ta_busy_per_cu": round(a["TA_TA_BUSY_sum"] / 256.0), "td_busy_per_cu": round(a["TD_TD_BUSY_sum"] / 256.0),
         "engine_clock_hz": round(clock) if clock else None,
         "kernel_ns_in_ta_pass": round(dur) if dur else None}
    with open(COUNTERS, "w") as fo:
        json.dump(c, fo, indent=1)
        fo.write("\n")
    return c


def _dispatch_ns(path, kernel):
    """Average duration (ns) of the kernel's dispatches in a counter pass,
    the first excluded."""
    import csv
    seen = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if kernel in r["Kernel_Name"]:
                seen[int(r["Dispatch_Id"])] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    v = [seen[k] for k in sorted(seen)][1:]
    return sum(v) / len(v) if v else None


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--summarize-profiles":
        print(json.dumps(summarize_profiles(sys.argv[2]), indent=1))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the headline trace (for profiling)")
    ap.add_argument("--extras", default="all",
                    help="comma list of the extra workloads to run: reference, backward, lights, reflection, "
                         "batched, prefabs, large (default all)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed region, copy 1 GiB device-to-device (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from voxelengine_amd import api, configs, views

    cfg = configs.CONFIGS["F"]
    W, H = cfg["width"], cfg["height"]
    params = configs.trace_params(**cfg["params"])
    ctx = api.Context(0)
    stream = torch.cuda.current_stream()

    # ---- scene: generate, mips, rebuild occupancy + levels (untimed setup, timed separately)
    nx, ny, nz = cfg["dims"]
    model = api.Model(ctx, nx, ny, nz)
    model.generate(seed=configs.SEED, terrain_fill=cfg["fill"], nboxes=cfg["nboxes"], box_min=cfg["box"][0],
                   box_max=cfg["box"][1], grounded=cfg["grounded"])
    model.build_mips()
    occ = api.Volume(ctx, nx // 2, ny // 2, nz // 2, nlevels=cfg["levels"])
    inst = [api.make_instance(model.mip_ptr(0), model.dims, api.IDENTITY, api.IDENTITY)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    occ.clear()
    occ.rebuild(inst)
    occ.build_levels()
    torch.cuda.synchronize()
    rebuild_ms = (time.perf_counter() - t0) * 1e3

    vd = views.scene_view(cfg["dims"], W, H)
    g = api.GBuffer(W, H)
    api.primary(ctx, model, vd, g)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        api.primary(ctx, model, vd, g)
    e1.record(stream)
    torch.cuda.synchronize()
    primary_ms = e0.elapsed_time(e1) / 3

    pal = configs.make_palette()
    pal_d = torch.from_numpy(pal).cuda()
    mat = g.material.cpu().numpy()
    hit_px = int((mat > 0).sum())
    rays_per_frame = hit_px * configs.rays_per_pixel(params)

    # algorithmic bytes: reference fetch count of every ray (one counting launch)
    vis = api.Visibility(W, H, fetches=True)
    api.trace(ctx, occ, vd, g, pal_d, params, vis)
    torch.cuda.synchronize()
    s_ref = int(vis.fetches.to(torch.int64).sum().item())
    alg_bytes = s_ref + PIXEL_IO_BYTES * hit_px
    # fetches the kernel executes: a ray that ends by the step limit or maxT is
    # marched once, where the reference repeats its pass four times
    # (Light.frag:79); from every ray's hit record (term, steps, fetches)
    hits = api.trace_hits_device(ctx, occ, vd, g, pal_d, params, api.Visibility(W, H))
    term, steps, fetches = hits[:, 0].to(torch.int64), hits[:, 1].to(torch.int64), hits[:, 2].to(torch.int64)
    s_exec = int(torch.where(term == 0, steps // 4, fetches).sum().item())
    del hits, term, steps, fetches
    torch.cuda.empty_cache()
    vis_t = api.Visibility(W, H, fetches=False)

    # ---- timed region: K trace launches on the current stream; the library
    # records HIP events around its march kernel on that same stream
    for _ in range(args.warmup):
        api.trace(ctx, occ, vd, g, pal_d, params, vis_t)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.profile(True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        api.trace(ctx, occ, vd, g, pal_d, params, vis_t)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    march_total_ms, launches = ctx.profile_read()
    ctx.profile(False)
    trace_ms = ev0.elapsed_time(ev1) / args.steps
    kernel_ms = march_total_ms / max(launches, 1)
    ms_per_step = wall * 1e3 / args.steps
    value = rays_per_frame / (ms_per_step * 1e-3) / 1e9
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    exec_bytes = s_exec + PIXEL_IO_BYTES * hit_px
    counters = profiled_counters()
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    out = {
        "metric": "Grays/s (shadow+AO+spec-occl) at 1080p/512^3",
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "F: 512^3 terrain+2000 boxes, 2 levels, 1920x1080, 1 shadow + 16 AO(32) + "
                               "1 GGX spec-occ ray per hit pixel, DDA marcher",
                   "hit_pixels": hit_px, "rays_per_frame": rays_per_frame, "marcher": "dda"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": counters["traffic_bytes_per_launch"] if counters else None,
                     "kernel": "k_march", "kernel_ms": round(kernel_ms, 4), "trace_ms": round(trace_ms, 4),
                     "alg_bytes_per_launch": alg_bytes, "ref_fetches_per_launch": s_ref,
                     "executed_fetches_per_launch": s_exec,
                     "frac_executed": round(exec_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "limiter": limiter(counters, kernel_ms)},
        "rebuild_ms": round(rebuild_ms, 3),
        "primary_ms": round(primary_ms, 4),
    }

    extras = set() if args.no_extras else (
        {"reference", "backward", "lights", "reflection", "batched", "prefabs", "large"} if args.extras == "all"
        else set(args.extras.split(",")))
    if "reference" in extras:
        # same frame through the reference's enabled marchers (Sparse / SuperSparse / Sparse)
        p_ref = configs.trace_params(**dict(cfg["params"], march="reference"))
        for _ in range(2):
            api.trace(ctx, occ, vd, g, pal_d, p_ref, vis_t)
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(10):
            api.trace(ctx, occ, vd, g, pal_d, p_ref, vis_t)
        ev1.record(stream)
        torch.cuda.synchronize()
        out["reference_marchers_grays_per_s"] = round(rays_per_frame / (ev0.elapsed_time(ev1) / 10 * 1e-3) / 1e9, 4)

    if "backward" in extras:
        # backward: shading adjoint over the same frame (visibility fixed)
        light = configs.default_light()
        rgb = torch.zeros((W * H, 3), dtype=torch.float32, device="cuda")
        api.shade_forward(ctx, vd, g, vis, pal_d, light, rgb)
        dl = torch.randn((W * H, 3), dtype=torch.float32, device="cuda")
        dpal = torch.zeros((256, 6), dtype=torch.float32, device="cuda")
        dlight = torch.zeros(6, dtype=torch.float32, device="cuda")
        for _ in range(3):
            api.shade_backward(ctx, vd, g, vis, pal_d, light, dl, dpal, dlight)
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(20):
            api.shade_backward(ctx, vd, g, vis, pal_d, light, dl, dpal, dlight)
        ev1.record(stream)
        torch.cuda.synchronize()
        bwd_ms = ev0.elapsed_time(ev1) / 20
        out["backward"] = {"ms": round(bwd_ms, 4), "grays_per_s": round(rays_per_frame / (bwd_ms * 1e-3) / 1e9, 2),
                           "px_per_s": round(hit_px / (bwd_ms * 1e-3), 1)}

    if "lights" in extras:
        log("light pass ...")
        out["lights"] = bench_lights(ctx, occ, vd, g, pal_d, cfg["dims"])
    if "reflection" in extras:
        out.update(bench_reflection_ssao(ctx, occ, vd, g, pal_d))
    if "batched" in extras:
        out["batched"] = bench_batched(ctx, pal_d)
    if "prefabs" in extras:
        log("prefabs ...")
        out["prefabs"] = bench_prefabs(ctx)
    if "large" in extras:
        log("config L ...")
        out["large_batched"] = bench_large(ctx, pal_d)

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        occ_h = occ.download()
        out["cpu_baseline"] = cpu_baseline(occ_h, vd, params, g.pos.cpu().numpy(), g.normal.cpu().numpy(), mat, pal)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
