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
CPU_SAMPLE_RADIUS = 131072
# HBM bytes per scan_band_kernel launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of
# the full-world run, scaled by the same counters on a copy of known size;
# profiles/scan_band_kernel_full_world_thr60.md): 165.4 KiB x 2.0 + 0.03 KiB x 1.0.
TRAFFIC_BYTES = 338800


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
    threads = max(1, min(16, os.cpu_count() or 1))
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
            "sample": sample, "seconds": dt, "gpu_matches_sample": gpu == result}


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
    peak = cus * 64 * clk_khz * 1e3 / 1e12  # INT32 VALU lane-ops/s in T
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
    roofline = {
        "bound": "valu", "unit": "TOP/s", "achieved": achieved, "peak": peak,
        "frac": achieved / peak,
        "traffic": args.traffic_kib * 1024 if args.traffic_kib is not None else TRAFFIC_BYTES,
        "kernel": dominant,
        "ops_per_center": OPS_PER_CENTER, "avg_launch_ms": d["avg_launch_ms"],
        "centers_per_launch": d["centers_per_launch"],
        "peak_basis": f"{cus} CUs x 64 lanes x {clk_khz / 1e6:.2f} GHz (INT32 VALU)",
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
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
