#!/usr/bin/env python3
"""Benchmark: full-world slime-cluster scan at threshold 60 on one MI355X.

Workload (one step = one complete search through the C ABI):
  seed = splitmix64 first output from state 1 (0x910A2DEC89025CC1),
  radius 1,875,000 chunks around (0, 0) -> 3,750,001^2 = 14,062,507,500,001
  candidate centers, minimum cluster size 60, top 65,536 rows kept.
The unit is T centers/s (10^12 candidate centers per second). Inputs are
synthetic (a seed); nothing is read from disk. Results are checked to be
identical across steps.

Printed as one JSON line with the kernel's INT32-VALU roofline (algorithmic
ops per center x centers / kernel time, kernel time from HIP events on the
launch stream) and a CPU baseline: the reference's SlimeCore search compiled
from its sources (oracle/_ref), timed on a bounded sample of the same
workload on this host's cores.

Two more legs ride in the same line:
  batched  256 searches from a splitmix64 stream (radius log-uniform in
           [256, 200000], centers within +-1.8M chunks, thresholds 50..62),
           run as one sweep (one launch per kernel class, optionally replayed
           from a hipGraph, or the persistent kernels) against 256 separate
           calls of the same library and 256 calls of its plain per-chunk
           kernel; all results must be identical. A secondary leg,
           batched_small, does the same for 256 searches of radius 256..4096.
           Their top sites (up to 1000 per search) then go through the biome
           filter (Deep Dark / Mushroom Fields under the sites' slime chunks,
           Minecraft 26.2) in one launch, against one context per world, and
           through the AFK-Y search (refine_candidates_y, heights -64..320
           step 4) in one launch, against one call per world and the
           reference's CPU refinement on a sample.
  served   Poisson arrivals of the served mix (70% radius 256..4096, 25%
           radius 50,000, 5% full world; thresholds 50..65; 10% pause then
           resume, 5% cancel) into one session, against the same arrivals
           served one search at a time in arrival order; p50/p99 latency and
           aggregate T/s of both, results identical. A secondary leg,
           served_small, serves 2000 requests of the small kind alone at
           200,000 arrivals per second, with no pause or cancel.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "T centers/s full-world thr60; batched+served vs plain aggregate T/s, p99 latency"
FULL_WORLD_SEED = -7995527694508729151  # splitmix64(1) = 0x910A2DEC89025CC1
FULL_WORLD_RADIUS = 1875000
THRESHOLD = 60
TOP_K = 65536
REFERENCE_TPS = 1.409  # published: 14.06 T centers in 9.979 s (README.md:15,19)
OPS_PER_CENTER = 6      # algorithmic INT32 ops per candidate center
# MI355X's published dense FP32 vector rate, 157.3 TFLOP/s, is 78.6 T lane
# operations per second (an FMA counts two flops): 256 CUs x 128 lanes x
# 2.4 GHz. The integer scan is read against it; single-cycle 32-bit VOP2
# operations (add, xor, and, shifts) issue at up to that rate, VOP3 forms
# (multiplies, bcnt, alignbit, three-operand adds) at half of it.
VALU_PEAK_T = 157.3 / 2
# VALU lane instructions per center of the full-world search's banded kernels,
# counted by rocprofv3 (SQ_INSTS_VALU x 64 / centers over all launches;
# profiles/scan_band_kernel_full_world_thr60.md).
VALU_INSTS_PER_CENTER = 4.59
CPU_SAMPLE_RADIUS = 131072
# HBM bytes per scan_band_kernel launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of
# the full-world run, scaled by the same counters on a copy of known size;
# profiles/scan_band_kernel_full_world_thr60.md): 81.4 KiB x 2.0 + 0.05 KiB x 1.0
# per launch, averaged over the search's launches after the first two (those
# two share their counter windows with the first-draw fix work running on the
# second stream).
TRAFFIC_BYTES = 166754


def host_cpu():
    """The host's CPU as the CPU baseline uses it: model, logical and physical
    cores, SMT, and the threads this process may run on (its CPU affinity,
    capped by a cgroup CPU quota where one is set: the box's CPU share)."""
    info = {"model": None, "logical_cores": os.cpu_count(), "physical_cores": None, "smt": None}
    try:
        cores, model = set(), None
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
        info["model"] = model
        info["physical_cores"] = len(cores) or None
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            info["smt"] = f.read().strip() == "1"
    except OSError:
        if info["physical_cores"]:
            info["smt"] = info["logical_cores"] > info["physical_cores"]
    usable = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    info["affinity_cpus"] = usable
    info["cgroup_cpu_quota"] = quota
    info["threads"] = min(usable, quota) if quota else usable
    return info


def hip_sync():
    try:
        C.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes: a kernel of known HBM traffic
    (read nbytes + write nbytes) that scales rocprofv3's FETCH_SIZE/WRITE_SIZE."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(gpu_lib):
    """Times the reference SlimeCore (or the scalar oracle port) on a sample."""
    from oracle import pyoracle as P
    import slime_amd
    cpu = host_cpu()
    threads = cpu["threads"]
    ref = P.ref_lib()
    rd = CPU_SAMPLE_RADIUS
    centers = (2 * rd + 1) ** 2
    if ref is not None:
        kind = "reference"
        P.ref_search(ref, FULL_WORLD_SEED, 1024, THRESHOLD, 221, 0, 0, 0, 1024, threads)  # warm-up
        times, result = [], None
        for _ in range(3):
            t0 = time.perf_counter()
            result = P.ref_search(ref, FULL_WORLD_SEED, rd, THRESHOLD, 221, 0, 0, 0, TOP_K, threads)
            times.append(time.perf_counter() - t0)
        dt = statistics.median(times)
        sample = (f"radius {rd} around (0,0) (= {centers} centers) of the full-world workload, "
                  f"same seed and threshold; median of 3 after a warm-up; "
                  f"{'AVX2' if P.cpu_has_avx2() else 'scalar'} build")
    else:
        kind, threads, rd = "port", 1, 4096
        centers = (2 * rd + 1) ** 2
        t0 = time.perf_counter()
        result = P.oracle_search(FULL_WORLD_SEED, rd, THRESHOLD, 221, 0, 0, 0, TOP_K)
        dt = time.perf_counter() - t0
        sample = f"radius {rd} around (0,0) (= {centers} centers), scalar C oracle, one run"
    gpu = slime_amd.search_slime_clusters_centered(FULL_WORLD_SEED, rd, THRESHOLD, 221, 0, 0, 0, TOP_K)
    return {"value": centers / dt / 1e12, "unit": "T centers/s", "cores": threads, "kind": kind,
            "sample": sample, "seconds": dt, "gpu_matches_sample": gpu == result, "host": cpu}


C1 = (0x123456789ABCDEF0, 2048, 40)  # seed, radius, threshold of the CPU-runnable search


def c1_leg():
    """The reference's CPU-runnable search (radius 2,048 around (0,0), threshold
    40, every hit returned): the GPU call against the reference SlimeCore on
    the host's cores, rows compared. Both are timed as C ABI calls into a
    result buffer allocated once (the latency a frontend sees); the Python
    wrapper's figure (a new 65,536-row buffer and tuple rows per call) is
    reported beside it."""
    from oracle import pyoracle as P
    import slime_amd
    seed, rd, thr = C1
    k = 65536
    centers = (2 * rd + 1) ** 2
    gpu = slime_amd.search_slime_clusters_centered(seed, rd, thr, 221, 0, 0, 0, k)
    lib = slime_amd.lib()
    buf = (slime_amd.ExtChunkResult * k)()
    times = []
    for _ in range(21):
        t0 = time.perf_counter()
        n = lib.search_slime_clusters_centered(seed, rd, thr, 221, 0, 0, 0, buf, k, 0, 0)
        times.append(time.perf_counter() - t0)
        if n != gpu[0]:
            raise SystemExit("C1 results differ between runs")
    wrapped = []
    for _ in range(5):
        t0 = time.perf_counter()
        got = slime_amd.search_slime_clusters_centered(seed, rd, thr, 221, 0, 0, 0, k)
        wrapped.append(time.perf_counter() - t0)
        if got != gpu:
            raise SystemExit("C1 results differ between runs")
    out = {"workload": f"seed {seed:#x}, radius {rd} around (0,0), threshold {thr}, all hits",
           "centers": centers, "hits": gpu[0], "gpu_ms": statistics.median(times) * 1e3,
           "gpu_ms_python_wrapper": statistics.median(wrapped) * 1e3,
           "gpu_t_centers_per_s": centers / statistics.median(times) / 1e12}
    ref = P.ref_lib()
    if ref is not None:
        threads = host_cpu()["threads"]
        want = P.ref_search(ref, seed, rd, thr, 221, 0, 0, 0, k, threads)  # warm-up, rows
        rbuf = (P.ChunkRow * k)()
        ct = []
        for _ in range(5):
            ref.reset_cancel()
            t0 = time.perf_counter()
            ref.search_slime_clusters_centered(seed, rd, thr, 221, 0, 0, 0, rbuf, k, threads, 0)
            ct.append(time.perf_counter() - t0)
        out.update({"reference_cpu_ms": statistics.median(ct) * 1e3, "reference_cores": threads,
                    "identical": want == gpu})
    print(f"C1: {gpu[0]} hits, gpu {out['gpu_ms']:.3f} ms", file=sys.stderr, flush=True)
    return out


def batched_leg(repeats):
    """The batched configuration: sweep vs separate calls (same results)."""
    import slime_amd
    from slime_amd.workloads import batched_searches, BATCH_STATE
    searches = batched_searches(256)
    cols = list(zip(*searches))
    k = 1000
    centers = sum((2 * s[1] + 1) ** 2 for s in searches)
    res = slime_amd.search_batch(*cols, max_results=k)  # warm-up, calibrates the autotuner
    out = {"workload": "256 searches, radius log-uniform [256, 200000], thr 50..62",
           "state": hex(BATCH_STATE), "centers": centers, "max_results": k}
    # one workgroup per tile (default), the same from a graph, and the
    # persistent work-queue kernels
    for key, graph, persistent in (("sweep", False, False), ("sweep_graph", True, False),
                                   ("sweep_persistent", False, True)):
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            got = slime_amd.search_batch(*cols, max_results=k, use_graph=graph,
                                         persistent=persistent)
            times.append(time.perf_counter() - t0)
            if got != res:
                raise SystemExit("batched results differ between runs")
        dt = min(times)
        st = slime_amd.batch_stats()
        out[key] = {"seconds": dt, "t_centers_per_s": centers / dt / 1e12,
                    "kernel_s": st["kernel_ns"] / 1e9, "launches": st["launches"],
                    "jobs": st["jobs"], "fallbacks": st["fallbacks"],
                    "hits_stored": st["hits_stored"], "hits_copied": st["hits_copied"]}
        print(f"batched {key}: {dt:.3f} s", file=sys.stderr, flush=True)
    # separate calls: a warm-up pass (the first search of 2^37+ centers on a
    # device measures the per-search autotuner's costs), then the same
    # repeats as the sweep, best time
    per = [slime_amd.search_slime_clusters_gpu_v1_centered(*s, max_results=k) for s in searches]
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        again = [slime_amd.search_slime_clusters_gpu_v1_centered(*s, max_results=k)
                 for s in searches]
        times.append(time.perf_counter() - t0)
        if again != per:
            raise SystemExit("per-search results differ between runs")
    dt = min(times)
    out["per_search"] = {"seconds": dt, "t_centers_per_s": centers / dt / 1e12}
    print(f"batched per-search: {dt:.3f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    plain = []
    for s in searches:
        seed, rd, thr, mx, _, cx, cz = s
        tot, rows, _ = slime_amd.scan_bands(seed, cx - rd, cz - rd, 2 * rd + 1, 2 * rd + 1, thr, mx,
                                            0, slime_amd.MODE_PLAIN, k)
        plain.append((tot, rows))
    dt = time.perf_counter() - t0
    out["per_search_plain_kernel"] = {"seconds": dt, "t_centers_per_s": centers / dt / 1e12}
    print(f"batched per-search plain kernel: {dt:.3f} s", file=sys.stderr, flush=True)
    best = max(out["sweep"]["t_centers_per_s"], out["sweep_graph"]["t_centers_per_s"])
    out["speedup_vs_per_search"] = best / out["per_search"]["t_centers_per_s"]
    out["speedup_vs_per_search_plain_kernel"] = best / out["per_search_plain_kernel"]["t_centers_per_s"]
    out["identical"] = per == res and plain == res
    out["biome"] = biome_leg(searches, res)
    out["afk_y"] = afk_y_leg(searches, res)
    return out


def batched_small_leg():
    """A batch of small searches only (radius log-uniform in [256, 4096], the
    same stream and thresholds): one sweep against separate calls, where the
    separate calls leave most of the device idle."""
    import slime_amd
    from slime_amd.workloads import batched_searches
    searches = batched_searches(256, r_hi=4096)
    cols = list(zip(*searches))
    k = 1000
    centers = sum((2 * s[1] + 1) ** 2 for s in searches)
    res = slime_amd.search_batch(*cols, max_results=k)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        got = slime_amd.search_batch(*cols, max_results=k)
        times.append(time.perf_counter() - t0)
        if got != res:
            raise SystemExit("batched results differ between runs")
    dt = min(times)
    st = slime_amd.batch_stats()
    t0 = time.perf_counter()
    per = [slime_amd.search_slime_clusters_gpu_v1_centered(*s, max_results=k) for s in searches]
    dp = time.perf_counter() - t0
    print(f"batched small: {dt:.4f} s, per-search {dp:.4f} s", file=sys.stderr, flush=True)
    return {"workload": "256 searches, radius log-uniform [256, 4096], thr 50..62",
            "centers": centers,
            "sweep": {"seconds": dt, "t_centers_per_s": centers / dt / 1e12,
                      "kernel_s": st["kernel_ns"] / 1e9, "prepare_s": st["prepare_ns"] / 1e9,
                      "sweep_s": st["sweep_ns"] / 1e9, "wall_s": st["wall_ns"] / 1e9},
            "per_search": {"seconds": dp, "t_centers_per_s": centers / dp / 1e12},
            "speedup_vs_per_search": dp / dt, "identical": per == res}


FP64_PEAK_TF = 78.6          # MI355X dense FP64 vector peak
FLOPS_PER_COLUMN = 52 * 60   # 52 double Perlin octave samples of ~60 FP64 ops
# FP64 add and multiply wave instructions of one filter launch over these
# sites, counted by rocprofv3 (SQ_INSTS_VALU_FLOPS_FP64, no FMA: the noise is
# compiled without contraction; profiles/biome_filter_batched_top_sites.md),
# times 64 lanes
FP64_FLOPS_COUNTED = 4.285e8 * 64
EXTENDED_LEVELS = (-64, 320, 4)  # block heights of the wider biome sweep


def biome_leg(searches, res):
    """The batched + biome configuration: every search's top sites (up to 1000)
    through the Deep Dark / Mushroom Fields filter of its own world, in one
    launch, against one context per world (the reference's way of calling)."""
    import slime_amd
    seeds = [s[0] for s in searches]
    cands = [[((r[1] - 8) // 16, (r[2] - 8) // 16) for r in rows] for _, rows in res]
    sizes = [r[0] for _, rows in res for r in rows]
    n = len(sizes)
    slime_amd.filter_multi(seeds, cands, 3)  # warm-up
    t0 = time.perf_counter()
    masks, ms = slime_amd.filter_multi(seeds, cands, 3)
    wall = time.perf_counter() - t0
    flat = [m for ws in masks for m in ws]
    t0 = time.perf_counter()
    per, per_ms = [], 0.0
    for seed, cs in zip(seeds, cands):
        if not cs:
            per.append([])
            continue
        with slime_amd.BiomeContext(seed, capacity=len(cs)) as ctx:
            per.append(ctx.filter([c[0] for c in cs], [c[1] for c in cs], 3))
            per_ms += ctx.last_filter_ms()
    per_wall = time.perf_counter() - t0
    columns = 16 * sum(sizes)  # quart columns under the sites' slime chunks (before early exits)
    achieved = FP64_FLOPS_COUNTED / (ms / 1e3) / 1e12
    print(f"biome filter: {n} sites in {ms:.1f} ms", file=sys.stderr, flush=True)
    # the wider sweep: both biomes at every height -64..320 step 4, over the
    # disk's slime chunks and over every chunk of the disk
    extended = {"heights": "-64..320 step 4 (quart y -16..80), both biomes at every height"}
    for name, flags in (("slime_chunks", 3), ("all_chunks", 3 | slime_amd.ALL_CHUNKS)):
        m_ext, ms_ext = slime_amd.filter_multi(seeds, cands, flags, levels=EXTENDED_LEVELS)
        fl = [m for ws in m_ext for m in ws]
        extended[name] = {"kernel_ms": ms_ext, "sites_per_s": n / (ms_ext / 1e3),
                          "deep_dark": sum(1 for m in fl if m & 1),
                          "mushroom_fields": sum(1 for m in fl if m & 2)}
        print(f"biome filter, heights -64..320, {name}: {ms_ext:.1f} ms", file=sys.stderr,
              flush=True)
    return {
        "workload": "top sites (<= 1000) of each batched search, flags Deep Dark | Mushroom Fields, "
                    "version 26.2",
        "sites": n, "deep_dark": sum(1 for m in flat if m & 1),
        "mushroom_fields": sum(1 for m in flat if m & 2),
        "one_launch": {"kernel_ms": ms, "wall_s": wall, "sites_per_s": n / (ms / 1e3)},
        "per_world_contexts": {"kernel_ms": per_ms, "wall_s": per_wall},
        "identical": per == masks,
        "extended": extended,
        "roofline": {"bound": "fp64", "unit": "TFLOP/s", "peak": FP64_PEAK_TF,
                     "achieved": achieved, "frac": achieved / FP64_PEAK_TF,
                     "flops_counted": FP64_FLOPS_COUNTED,
                     "flops_modelled": columns * FLOPS_PER_COLUMN, "columns": columns,
                     "note": "FP64 is 17% of the VALU instructions; the parameter-tree "
                             "searches (integer, latency-bound LDS reads) are most of the "
                             "time"},
    }


AFK_Y = (1, -64, -64, 320, 4)   # y scan on, platform -64, heights -64..320 step 4
REFINE_VALU_PER_COLUMN = 22     # VALU instructions per block column in refine_kernel's loops


def refine_columns_per_point(enabled, platform_y, y_min, y_max, y_step):
    """Block columns one AFK point walks over all heights within reach:
    2*reach+1 of the outer disk plus 2*reach_in+1 of the inner one."""
    import math
    y_step = y_step if y_step > 0 else 4
    lo, hi = (min(y_min, y_max), max(y_min, y_max)) if enabled else (platform_y, platform_y)
    step = y_step if enabled else 1
    cols = 0
    for y in range(lo, hi + 1, step):
        dy2 = (platform_y - y) ** 2
        if dy2 <= 16384:
            cols += 2 * math.isqrt(16384 - dy2) + 1
            if dy2 <= 576:
                cols += 2 * math.isqrt(576 - dy2) + 1
    return cols


def afk_y_leg(searches, res):
    """The AFK-Y search over the batched top sites: refine_candidates_y
    (81 AFK points x heights -64..320 step 4 around platform -64) of every
    world's rows in one launch, against one call per world, and against the
    reference's CPU refine_candidates_y on a sample of worlds."""
    import slime_amd
    seeds = [s[0] for s in searches]
    rows = [list(r) for _, r in res]
    n = sum(len(r) for r in rows)
    slime_amd.set_y_scan_config(*AFK_Y)
    try:
        slime_amd.refine_multi(seeds, rows)  # warm-up
        t0 = time.perf_counter()
        got = slime_amd.refine_multi(seeds, rows)
        wall = time.perf_counter() - t0
        ns, evals = slime_amd.refine_stats()
        t0 = time.perf_counter()
        per, per_ns = [], 0
        for seed, rs in zip(seeds, rows):
            per.append(slime_amd.refine_candidates_y(seed, rs) if rs else [])
            per_ns += slime_amd.refine_stats()[0] if rs else 0
        per_wall = time.perf_counter() - t0
        cpu = afk_y_cpu(seeds, rows, got)
    finally:
        slime_amd.set_y_scan_config(0, -64, -64, 64, 4)
    peak = VALU_PEAK_T
    columns = n * 81 * refine_columns_per_point(*AFK_Y)
    achieved = columns * REFINE_VALU_PER_COLUMN / (ns / 1e9) / 1e12
    print(f"afk-y refine: {n} rows in {ns / 1e6:.2f} ms", file=sys.stderr, flush=True)
    return {
        "workload": "refine_candidates_y of the batched top sites, y scan -64..320 step 4, "
                    "platform -64, 81 AFK points",
        "rows": n, "evaluations": evals,
        "one_launch": {"kernel_ms": ns / 1e6, "wall_s": wall, "rows_per_s": n / (ns / 1e9)},
        "per_world_calls": {"kernel_ms": per_ns / 1e6, "wall_s": per_wall},
        "identical": per == got,
        "cpu_baseline": cpu,
        "roofline": {"bound": "valu", "unit": "TOP/s", "peak": peak, "achieved": achieved,
                     "frac": achieved / peak, "columns": columns,
                     "valu_per_column": REFINE_VALU_PER_COLUMN,
                     "note": "issue rate of the compiled loop (VALU instructions per block "
                             "column x columns / kernel time)"},
    }


def afk_y_cpu(seeds, rows, got):
    """The compiled reference's refine_candidates_y over the first worlds'
    rows (about 2,000), on this host's cores."""
    from oracle import pyoracle as P
    ref = P.ref_lib()
    if ref is None:
        return None
    threads = host_cpu()["threads"]
    sample, same, t = 0, True, 0.0
    for w, (seed, rs) in enumerate(zip(seeds, rows)):
        if sample >= 2000:
            break
        if not rs:
            continue
        t0 = time.perf_counter()
        want = P.ref_refine(ref, "refine_y", seed, rs, AFK_Y, threads=threads)
        t += time.perf_counter() - t0
        # The CPU library keeps the center for a row whose counts are all 0
        # (the GPU library takes point 0): such rows never rank here.
        same = same and want == got[w]
        sample += len(rs)
    ref.set_y_scan_config(0, -64, -64, 64, 4)
    return {"value": sample / t, "unit": "rows/s", "cores": threads, "kind": "reference",
            "sample": f"{sample} rows of the first worlds, same heights", "seconds": t,
            "gpu_matches_sample": same}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def latency_stats(lat, reqs, idx):
    """p50/p99 of the latencies of idx, overall and for the small requests
    (radius <= 4096, 70% of the mix)."""
    small = [lat[i] for i in idx if reqs[i]["search"][1] <= 4096]
    every = [lat[i] for i in idx]
    return {"p50_s": percentile(every, 50), "p99_s": percentile(every, 99),
            "small_p50_s": percentile(small, 50), "small_p99_s": percentile(small, 99)}


# Device time per served round. Rounds are pipelined (the next one is queued
# while the current one runs), so short rounds cost little throughput: on the
# served mix 10 ms rounds keep small requests near a p50 of 25-30 ms at about
# the aggregate of one-at-a-time calls.
SERVED_ROUND_MS = 10


def served_leg(n, rate, with_plain=False, small_only=False):
    """The served configuration: one session vs one search at a time through
    the same library's per-search engine (optionally also through the plain
    per-chunk kernel). At the default rate requests arrive faster than the
    device finishes them, so both forms are device-bound. small_only: every
    request of the mix's small kind (radius 256..4,096)."""
    import threading
    import slime_amd
    from slime_amd.workloads import served_requests, SERVED_STATE
    reqs = served_requests(n, rate, small_only=small_only)
    if small_only:  # batching alone: no pauses or cancels
        for r in reqs:
            r["action"] = None
    k = 1000
    pause_s, cancel_s = 0.005, 0.005
    # served: continuous batching
    lat, results = {}, {}
    with slime_amd.Session(hbm_budget=16 << 30, max_active=256 if small_only else 64,
                           round_ms=SERVED_ROUND_MS) as sess:
        ids = [None] * n
        t_start = time.perf_counter()

        def arrivals():
            for i, r in enumerate(reqs):
                delay = t_start + r["arrival"] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids[i] = sess.submit(*r["search"], max_results=k)
                if r["action"] == "pause":
                    threading.Timer(pause_s, sess.pause, (ids[i],)).start()
                    threading.Timer(pause_s + 0.05, sess.resume, (ids[i],)).start()
                elif r["action"] == "cancel":
                    threading.Timer(cancel_s, sess.cancel, (ids[i],)).start()

        th = threading.Thread(target=arrivals)
        th.start()
        th.join()
        t_last = t_start
        for i, r in enumerate(reqs):
            st = sess.wait(ids[i], -1)
            if st == slime_amd.DONE:
                state, _, total, lat_ns = sess.poll(ids[i])
                if r["action"] != "cancel":
                    lat[i] = lat_ns / 1e9
                    results[i] = sess.result(ids[i], k)
        t_last = time.perf_counter()
        stats = sess.stats()
    served_wall = t_last - t_start
    done = [i for i in results]
    centers = sum((2 * reqs[i]["search"][1] + 1) ** 2 for i in done)
    print(f"served: {len(done)} requests in {served_wall:.2f} s", file=sys.stderr, flush=True)
    def sequential(run, name):
        """One search at a time in arrival order (cancelled ones skipped, a
        paused one holds the device for the pause time)."""
        seq_lat, seq_res = {}, {}
        slime_amd.kernel_stats(reset=True)
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            now = time.perf_counter() - t0
            if now < r["arrival"]:
                time.sleep(r["arrival"] - now)
            if r["action"] == "cancel":
                continue
            if r["action"] == "pause":
                time.sleep(0.05)
            seq_res[i] = run(r["search"])
            seq_lat[i] = time.perf_counter() - t0 - r["arrival"]
        seq_wall = time.perf_counter() - t0
        _, kns, _ = slime_amd.kernel_stats()
        print(f"{name}: {len(seq_res)} requests in {seq_wall:.2f} s", file=sys.stderr, flush=True)
        seq_centers = sum((2 * reqs[i]["search"][1] + 1) ** 2 for i in seq_res)
        ident = all(results[i] == seq_res.get(i) for i in done) and set(done) == set(seq_res)
        out = {"wall_s": seq_wall, "t_centers_per_s": seq_centers / seq_wall / 1e12,
               "kernel_s": kns / 1e9}
        out.update(latency_stats(seq_lat, reqs, list(seq_res)))
        return out, ident

    seq, ident = sequential(
        lambda s: slime_amd.search_slime_clusters_gpu_v1_centered(*s, max_results=k), "sequential")

    def plain_search(s):
        seed, rd, thr, mx, _, cx, cz = s
        tot, rows, _ = slime_amd.scan_bands(seed, cx - rd, cz - rd, 2 * rd + 1, 2 * rd + 1, thr,
                                            mx, 0, slime_amd.MODE_PLAIN, k)
        return tot, rows

    seq_plain, ident_plain = (None, True) if not with_plain else sequential(
        plain_search, "sequential plain kernel")
    served = {"wall_s": served_wall, "t_centers_per_s": centers / served_wall / 1e12,
              "kernel_s": stats["kernel_ns"] / 1e9, "rounds": stats["rounds"]}
    served.update(latency_stats(lat, reqs, list(lat)))
    out = {
        "workload": ("Poisson arrivals: radius 256..4096 only (the served mix's small kind); "
                     "thr 50..65; no pause or cancel") if small_only else
                    "Poisson arrivals: 70% radius 256..4096, 25% radius 50000, 5% full world; "
                    "thr 50..65; 10% pause+resume, 5% cancel",
        "state": hex(SERVED_STATE), "requests": n, "rate_per_s": rate,
        "round_ms": SERVED_ROUND_MS, "completed": len(done),
        "centers": centers, "last_arrival_s": reqs[-1]["arrival"],
        "device_bound": reqs[-1]["arrival"] < 0.5 * served_wall,
        "served": served,
        "sequential": seq,
        "speedup_vs_sequential": served["t_centers_per_s"] / seq["t_centers_per_s"],
        "identical": ident and ident_plain,
    }
    if seq_plain is not None:
        # the same arrivals through the plain per-chunk kernel, one at a time
        out["sequential_plain_kernel"] = seq_plain
        out["speedup_vs_sequential_plain_kernel"] = (served["t_centers_per_s"] /
                                                     seq_plain["t_centers_per_s"])
    return out


def band_classes_leg(reps=7):
    """Kernel rate of each banded class with two x-phases and the union
    bound: strided sets of nb = 2..14 bands of 16,384 rows over 4 x-blocks
    (1,048,576 columns), threshold 55, seed 0x1234567; T centers per second
    of kernel time (median of `reps` scans)."""
    import statistics
    import slime_amd
    p = 1 << 17
    out = {}
    # one launch per class: no measured regrouping (the classes above 12 bands
    # also need SLIME_MAX_BANDS=14 in the environment)
    os.environ["SLIME_BAND_TUNE"] = "0"
    for nb in range(2, 15):
        w, rows = 8 * p, 16384
        args = (0x1234567, -3 * p, -2 * p, w, (nb - 1) * p + rows, 55, 221, rows,
                slime_amd.MODE_BANDED_UNION, 4096)
        slime_amd.scan_bands(*args)
        rates = []
        for _ in range(reps):
            slime_amd.kernel_stats(slime_amd.KERNEL_BANDED, reset=True)
            if slime_amd.scan_bands(*args)[0] < 0:
                raise SystemExit("scan_bands failed")
            _, ns, centers = slime_amd.kernel_stats(slime_amd.KERNEL_BANDED)
            rates.append(centers / ns / 1e3)
        out[str(nb)] = round(statistics.median(rates), 4)
        print(f"band class {nb}: {out[str(nb)]} T centers/s", file=sys.stderr, flush=True)
    return out


def window_ab_leg(blocks=3072, steps=4096, min_size=60, seed=0x5EED):
    """Counting stage alone (slime_window_ab): VALU word counting (method 0)
    against int8 MFMA window sums (method 1) over the same hashed bitplane
    ring, for 8 bands (the full-world launches) and 4 (the largest searches
    of the batched set). Rate: window centers counted per second, 4 rows x
    496 centers x 2 phases x bands per strip and workgroup; best of two
    launches; both methods' checksums must agree."""
    import slime_amd
    out = {}
    for bands in (8, 4):
        row = {}
        for method, name in ((slime_amd.WINDOW_VALU, "valu"), (slime_amd.WINDOW_MFMA, "mfma")):
            best, res = None, None
            for _ in range(2):
                r = slime_amd.window_ab(method, bands, blocks, steps, min_size, seed)
                if r is None:
                    raise SystemExit("window_ab failed")
                if res is not None and r[:2] != res[:2]:
                    raise SystemExit("window_ab: launches disagree")
                res = r
                best = r[2] if best is None else min(best, r[2])
            centers = blocks * steps * 4 * 496 * 2 * bands
            row[name] = {"kernel_ms": best / 1e6, "t_centers_per_s": centers / (best * 1e-9) / 1e12,
                         "checksum": res[0], "passes": res[1]}
            print(f"window_ab bands {bands} {name}: {best / 1e6:.3f} ms", file=sys.stderr,
                  flush=True)
        if row["valu"]["checksum"] != row["mfma"]["checksum"] or \
                row["valu"]["passes"] != row["mfma"]["passes"]:
            raise SystemExit("window_ab: methods disagree")
        row["mfma_over_valu_time"] = row["mfma"]["kernel_ms"] / row["valu"]["kernel_ms"]
        out[f"bands_{bands}"] = row
    out["workload"] = {"blocks": blocks, "steps": steps, "min_size": min_size, "seed": seed}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--radius", type=int, default=FULL_WORLD_RADIUS,
                    help="search radius in chunks (default: the full world)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device-to-device copy of BYTES (profiler scale)")
    ap.add_argument("--traffic-kib", type=float, default=None,
                    help="HBM KiB per launch of the dominant kernel measured by rocprofv3")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched leg (for profiling its kernels)")
    ap.add_argument("--only-biome", action="store_true",
                    help="run only the batched searches once and the biome filter of their top "
                         "sites (for profiling the filter)")
    ap.add_argument("--only-valu-rates", action="store_true",
                    help="run only the VALU issue-rate microbenchmark (for profiling it)")
    ap.add_argument("--only-band-classes", action="store_true",
                    help="run only the per-class rate of the banded kernels (two x-phases, "
                         "union bound, 2-14 bands)")
    ap.add_argument("--only-window-ab", action="store_true",
                    help="run only the counting-stage A/B (VALU word counting vs int8 MFMA "
                         "window sums) at 8 and 4 bands")
    ap.add_argument("--no-served", action="store_true", help="skip the served leg")
    ap.add_argument("--served-plain", action="store_true",
                    help="also serve the arrivals one at a time through the plain per-chunk "
                         "kernel (a slow secondary baseline)")
    ap.add_argument("--served-requests", type=int, default=120)
    ap.add_argument("--served-rate", type=float, default=60.0,
                    help="arrivals per second (the default outpaces the device)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the search shards by z-slabs across GPUs only in design).", file=sys.stderr)
        return 2

    import slime_amd
    lib = slime_amd.lib()
    if lib.get_cuda_device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    if args.only_batched:
        print(json.dumps({"batched": batched_leg(1)}), flush=True)
        return 0
    if args.only_biome:
        from slime_amd.workloads import batched_searches
        searches = batched_searches(256)
        res = slime_amd.search_batch(*zip(*searches), max_results=1000)
        print(json.dumps({"biome": biome_leg(searches, res)}), flush=True)
        return 0
    if args.only_valu_rates:
        rates = {name: slime_amd.valu_rate(kind, 16384) / 1e12
                 for kind, name in enumerate(slime_amd.VALU_KINDS)}
        print(json.dumps({"valu_rates_t_per_s": rates, "peak": VALU_PEAK_T}), flush=True)
        return 0
    if args.only_band_classes:
        print(json.dumps({"band_classes": band_classes_leg()}), flush=True)
        return 0
    if args.only_window_ab:
        print(json.dumps({"window_ab": window_ab_leg()}), flush=True)
        return 0
    rd = args.radius
    centers = (2 * rd + 1) ** 2
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    def step():
        return slime_amd.search_slime_clusters_centered(FULL_WORLD_SEED, rd, THRESHOLD, 221, 0,
                                                        0, 0, TOP_K)

    first = None
    for i in range(args.warmup):
        first = step()
        print(f"warmup {i}: {first[0]} hits", file=sys.stderr, flush=True)
    hip_sync()
    slime_amd.kernel_stats(reset=True)
    times = []
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = step()
        hip_sync()
        times.append(time.perf_counter() - t0)
        print(f"step {i}: {res[0]} hits in {times[-1]:.3f} s", file=sys.stderr, flush=True)
        if res[0] < 0:
            raise SystemExit("search failed")
        if first is not None and res != first:
            raise SystemExit("results differ between steps")
        first = res
    elapsed = time.perf_counter() - t_all
    ms_per_step = elapsed / args.steps * 1e3
    value = centers * args.steps / elapsed / 1e12

    info = slime_amd.device_info()
    cus, clk_khz = (info[0], info[1]) if info else (256, 2400000)
    peak = VALU_PEAK_T
    kernels = {}
    for kind, name in ((slime_amd.KERNEL_BANDED, "scan_band_kernel"),
                       (slime_amd.KERNEL_PLAIN, "scan_plain_kernel")):
        n, ns, c = slime_amd.kernel_stats(kind)
        if n:
            kernels[name] = {"launches": n, "kernel_s": ns / 1e9, "centers": c,
                             "avg_launch_ms": ns / 1e6 / n, "centers_per_launch": c / n,
                             "t_centers_per_s": c / (ns / 1e9) / 1e12}
    dominant = max(kernels, key=lambda k: kernels[k]["kernel_s"])
    d = kernels[dominant]
    achieved = d["centers"] * OPS_PER_CENTER / d["kernel_s"] / 1e12
    # the device's measured issue rate per instruction kind (T lane-instr/s)
    issue = {}
    for kind, name in enumerate(slime_amd.VALU_KINDS):
        r = slime_amd.valu_rate(kind, 16384)
        issue[name] = r / 1e12 if r else None
    valu_rate = VALU_INSTS_PER_CENTER * d["centers"] / d["kernel_s"] / 1e12
    roofline = {
        "bound": "valu", "unit": "TOP/s", "achieved": achieved, "peak": peak,
        "frac": achieved / peak,
        "traffic": args.traffic_kib * 1024 if args.traffic_kib is not None else TRAFFIC_BYTES,
        "kernel": dominant,
        "ops_per_center": OPS_PER_CENTER, "avg_launch_ms": d["avg_launch_ms"],
        "centers_per_launch": d["centers_per_launch"],
        "peak_basis": "published dense FP32 vector rate 157.3 TFLOP/s = 78.6 T lane-ops/s "
                      f"({cus} CUs x 128 lanes x 2.4 GHz)",
        "measured_issue_t_per_s": issue,
        "valu_insts_per_center": VALU_INSTS_PER_CENTER,
        "valu_insts_t_per_s": valu_rate,
        "valu_frac": valu_rate / peak,
        "kernels": kernels,
    }
    out = {
        "metric": METRIC, "value": value, "unit": "T centers/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": value / REFERENCE_TPS if rd == FULL_WORLD_RADIUS else None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "full_world_thr60", "seed": FULL_WORLD_SEED, "rd_max": rd,
                   "center": [0, 0], "min_size": THRESHOLD, "max_size": 221, "max_results": TOP_K,
                   "centers": centers},
        "hits": first[0], "top": first[1][:3],
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(lib)
        out["c1"] = c1_leg()
    if not args.no_batched:
        out["batched"] = batched_leg(2)
        # secondary: where separate calls cannot fill the device
        out["batched_small"] = batched_small_leg()
    if not args.no_served:
        out["served"] = served_leg(args.served_requests, args.served_rate, args.served_plain)
        # secondary: the mix's small requests alone, arriving faster than
        # either form serves them
        out["served_small"] = served_leg(2000, 200000.0, small_only=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
