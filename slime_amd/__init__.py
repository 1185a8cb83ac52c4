"""slime_amd: MI355X-native slime-chunk cluster search.

Python side of the drop-in: ctypes bindings to libslimefinder.so with the
reference's function names and argument meaning (SlimeCore.cpp:340-651,
SlimeCoreGPU.cu:3034-3955, called from Python via ctypes exactly as the
reference's GUI does: xp_scan.py:537-551). There is no Python or CPU
fallback: importing a function fails loudly if the native library is absent.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslimefinder.so")

__all__ = [
    "ExtChunkResult", "lib", "build", "search_slime_clusters_centered",
    "search_slime_clusters", "search_slime_clusters_gpu_v1_centered", "scan_rect",
    "EXPORTS", "set_y_scan_config", "score_candidates_precise", "refine_candidates_y",
    "refine_stats", "refine_multi",
]

# Every symbol include/slimefinder.h declares.
EXPORTS = (
    "search_slime_clusters_centered", "search_slime_clusters",
    "search_slime_clusters_gpu_v1_centered", "search_slime_clusters_gpu_v1",
    "search_slime_clusters_gpu_v34", "search_slime_clusters_gpu_v34_centered",
    "request_cancel", "request_pause", "resume_search", "reset_cancel", "get_progress",
    "get_processed_centers", "get_gpu_scan_work_ns", "set_y_scan_config", "is_slime_chunk_c",
    "get_cuda_device_count", "get_cuda_device_name", "cleanup_gpu_resources",
    "get_gpu_v1_shape", "get_gpu_v1_algorithm", "get_gpu_v1_rng", "set_gpu_v1_rng_override",
    "get_gpu_v34_shape", "get_gpu_v34_rng", "set_gpu_v34_rng_override", "slime_scan_rect",
    "slime_kernel_stats", "slime_device_info", "slime_scan_bands", "slime_search_batch",
    "slime_batch_stats", "slime_sweep_costs", "slime_session_create", "slime_session_destroy",
    "slime_session_submit", "slime_session_pause", "slime_session_resume",
    "slime_session_cancel", "slime_session_poll", "slime_session_wait", "slime_session_result",
    "slime_session_release", "slime_session_stats", "score_candidates_precise",
    "refine_candidates_y", "slime_refine_stats", "refine_candidates_y_multi",
    "slime_valu_rate", "slime_session_create_ex", "refine_candidates_y_noalloc",
    "slime_window_ab", "slime_batch_class_centers",
)

# Every symbol include/mcgpu.h declares (the biome filter).
BIOME_EXPORTS = (
    "mcgpu_api_version", "mcgpu_build_name", "mcgpu_modern_ctx_create",
    "mcgpu_modern_ctx_create_ex", "mcgpu_modern_ctx_destroy", "mcgpu_modern_ctx_filter_candidates",
    "mcgpu_modern_ctx_query", "mcgpu_modern_recommended_filter_batch", "mcgpu_modern_ctx_columns",
    "mcgpu_modern_ctx_last_filter_ms", "mcgpu_filter_multi", "mcgpu_filter_multi_last_ms",
    "mcgpu_xoroshiro_nextlong_batch", "mcgpu_xperlin_sample_batch",
    "mcgpu_xperlin_sample_fixed_seed", "mcgpu_continentalness_noise_fixed_seed",
    "mcgpu_climate5_fixed_seed", "mcgpu_climate5_quant_fixed_seed",
    "mcgpu_climate5_i32_fixed_seed", "mcgpu_climate_ctx_create", "mcgpu_climate_ctx_query_i32",
    "mcgpu_climate_ctx_destroy", "mcgpu_climate6_i32_fixed_seed",
    "mcgpu_climate6_i32_fixed_seed_ex", "mcgpu_biome20_i32_fixed_seed", "mcgpu_biome20_ctx_create",
    "mcgpu_biome20_ctx_query", "mcgpu_biome20_ctx_destroy",
    "mcgpu_modern_ctx_filter_candidates_levels", "mcgpu_filter_multi_levels",
)


class ExtChunkResult(C.Structure):
    """24-byte result row (SlimeCore.cpp:182-189)."""
    _fields_ = [("size", C.c_int32), ("center_x", C.c_int32), ("center_z", C.c_int32),
                ("obs_count", C.c_int32), ("afk_x", C.c_int32), ("afk_z", C.c_int32)]


def build(jobs=8):
    """Compiles libslimefinder.so for gfx950 with hipcc (no GPU needed)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"], check=True)


_lib = None


def lib():
    """The loaded native library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with slime_amd.build() "
                           "(there is no fallback path)")
    L = C.CDLL(LIB_PATH)
    i32, i64, P = C.c_int32, C.c_int64, C.POINTER(ExtChunkResult)
    L.search_slime_clusters_centered.argtypes = [i64, i64, i32, i32, i64, i32, i32, P, i32, i32, i32]
    L.search_slime_clusters_centered.restype = i64
    L.search_slime_clusters.argtypes = [i64, i64, i32, i32, i64, P, i32, i32, i32]
    L.search_slime_clusters.restype = i64
    for name in ("search_slime_clusters_gpu_v1_centered", "search_slime_clusters_gpu_v34_centered"):
        getattr(L, name).argtypes = [i64, i64, i32, i32, i64, i32, i32, P, i32, i32]
        getattr(L, name).restype = i64
    for name in ("search_slime_clusters_gpu_v1", "search_slime_clusters_gpu_v34"):
        getattr(L, name).argtypes = [i64, i64, i32, i32, i64, P, i32, i32]
        getattr(L, name).restype = i64
    L.slime_scan_rect.argtypes = [i64, i32, i32, i64, i64, i32, i32, P, i32, C.POINTER(C.c_int64)]
    L.slime_scan_rect.restype = i64
    for name in ("get_progress", "get_cuda_device_count", "get_gpu_v1_shape",
                 "get_gpu_v1_algorithm", "get_gpu_v1_rng", "get_gpu_v34_shape", "get_gpu_v34_rng"):
        getattr(L, name).restype = i32
        getattr(L, name).argtypes = []
    for name in ("get_processed_centers", "get_gpu_scan_work_ns"):
        getattr(L, name).restype = i64
        getattr(L, name).argtypes = []
    for name in ("request_cancel", "request_pause", "resume_search", "reset_cancel",
                 "cleanup_gpu_resources"):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = []
    L.set_y_scan_config.argtypes = [i32, i32, i32, i32, i32]
    L.set_y_scan_config.restype = None
    L.score_candidates_precise.argtypes = [i64, P, i32, i32, i32]
    L.score_candidates_precise.restype = None
    L.refine_candidates_y.argtypes = [i64, P, i32, i32]
    L.refine_candidates_y.restype = None
    L.refine_candidates_y_noalloc.argtypes = [i64, C.POINTER(ExtChunkResult), i32]
    L.refine_candidates_y_noalloc.restype = i32
    L.refine_candidates_y_multi.argtypes = [i32, C.POINTER(i64), C.POINTER(i32), P]
    L.refine_candidates_y_multi.restype = i32
    L.slime_refine_stats.argtypes = [C.POINTER(i64), C.POINTER(i64)]
    L.slime_refine_stats.restype = None
    L.is_slime_chunk_c.argtypes = [i64, i64, i64]
    L.is_slime_chunk_c.restype = C.c_bool
    L.get_cuda_device_name.argtypes = [C.c_char_p, i32]
    L.get_cuda_device_name.restype = i32
    L.set_gpu_v1_rng_override.argtypes = [i32]
    L.set_gpu_v34_rng_override.argtypes = [i32]
    L.slime_kernel_stats.argtypes = [i32, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), i32]
    L.slime_kernel_stats.restype = None
    L.slime_device_info.argtypes = [C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)]
    L.slime_device_info.restype = i32
    L.slime_valu_rate.argtypes = [i32, i32, C.POINTER(i64), C.POINTER(i64)]
    L.slime_valu_rate.restype = i32
    L.slime_window_ab.argtypes = [i32, i32, i32, i32, i32, C.c_uint32, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint64), C.POINTER(i64)]
    L.slime_window_ab.restype = i32
    L.slime_scan_bands.argtypes = [i64, i32, i32, i64, i64, i32, i32, i32, i32, P, i32,
                                   C.POINTER(C.c_int64)]
    L.slime_scan_bands.restype = i64
    PI64, PI32, PD = C.POINTER(i64), C.POINTER(i32), C.POINTER(C.c_double)
    L.slime_search_batch.argtypes = [i32, PI64, PI64, PI32, PI32, PI64, PI32, PI32, P, i32, PI64,
                                     i32]
    L.slime_search_batch.restype = i32
    L.slime_batch_stats.argtypes = [PI64, PI64, PI64, PI64, PI64, PI32, PI32, PI32, PI64, PI64]
    L.slime_batch_stats.restype = None
    L.slime_sweep_costs.argtypes = [PD, PD, PI32, PI32, i32]
    L.slime_sweep_costs.restype = i32
    L.slime_batch_class_centers.argtypes = [PI64, i32]
    L.slime_batch_class_centers.restype = i32
    VP = C.c_void_p
    L.slime_session_create.argtypes = [i64, i32, i32]
    L.slime_session_create.restype = VP
    L.slime_session_create_ex.argtypes = [i64, i32, i32, i32]
    L.slime_session_create_ex.restype = VP
    L.slime_session_destroy.argtypes = [VP]
    L.slime_session_destroy.restype = None
    L.slime_session_submit.argtypes = [VP, i64, i64, i32, i32, i64, i32, i32, i32]
    L.slime_session_submit.restype = i64
    for name in ("slime_session_pause", "slime_session_resume", "slime_session_cancel",
                 "slime_session_release"):
        getattr(L, name).argtypes = [VP, i64]
        getattr(L, name).restype = i32
    L.slime_session_poll.argtypes = [VP, i64, PI32, PI64, PI64]
    L.slime_session_poll.restype = i32
    L.slime_session_wait.argtypes = [VP, i64, i32]
    L.slime_session_wait.restype = i32
    L.slime_session_result.argtypes = [VP, i64, P, i32]
    L.slime_session_result.restype = i64
    L.slime_session_stats.argtypes = [VP, PI64, PI64, PI32, PI32, PI32, PI64, PI64]
    L.slime_session_stats.restype = None
    L.mcgpu_api_version.restype = C.c_int
    L.mcgpu_build_name.restype = C.c_char_p
    L.mcgpu_modern_ctx_create.argtypes = [i32, C.c_uint64, i32]
    L.mcgpu_modern_ctx_create.restype = VP
    L.mcgpu_modern_ctx_create_ex.argtypes = [i32, C.c_uint64, i32, i32]
    L.mcgpu_modern_ctx_create_ex.restype = VP
    L.mcgpu_modern_ctx_destroy.argtypes = [VP]
    L.mcgpu_modern_ctx_destroy.restype = None
    L.mcgpu_modern_ctx_filter_candidates.argtypes = [VP, PI32, PI32, i32, i32, PI32]
    L.mcgpu_modern_ctx_filter_candidates.restype = C.c_int
    L.mcgpu_modern_ctx_query.argtypes = [VP, PI32, PI32, PI32, i32, PI32]
    L.mcgpu_modern_ctx_query.restype = C.c_int
    L.mcgpu_modern_ctx_columns.argtypes = [VP, PI32, PI32, i32, PI32]
    L.mcgpu_modern_ctx_columns.restype = C.c_int
    L.mcgpu_modern_recommended_filter_batch.restype = i32
    L.mcgpu_modern_ctx_last_filter_ms.argtypes = [VP]
    L.mcgpu_modern_ctx_last_filter_ms.restype = C.c_float
    U64, PU64, PD = C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_double)
    L.mcgpu_filter_multi.argtypes = [i32, i32, PU64, PI32, PI32, PI32, i32, i32, PI32]
    L.mcgpu_filter_multi.restype = C.c_int
    L.mcgpu_filter_multi_last_ms.restype = C.c_float
    L.mcgpu_modern_ctx_filter_candidates_levels.argtypes = [VP, PI32, PI32, i32, i32, i32, i32,
                                                            i32, PI32]
    L.mcgpu_modern_ctx_filter_candidates_levels.restype = C.c_int
    L.mcgpu_filter_multi_levels.argtypes = [i32, i32, PU64, PI32, PI32, PI32, i32, i32, i32, i32,
                                            i32, PI32]
    L.mcgpu_filter_multi_levels.restype = C.c_int
    L.mcgpu_xoroshiro_nextlong_batch.argtypes = [PU64, i32, i32, PU64]
    L.mcgpu_xperlin_sample_batch.argtypes = [PU64, PD, PD, PD, i32, PD]
    L.mcgpu_xperlin_sample_fixed_seed.argtypes = [U64, PD, PD, PD, i32, PD]
    L.mcgpu_continentalness_noise_fixed_seed.argtypes = [U64, PD, PD, i32, PD]
    L.mcgpu_climate5_fixed_seed.argtypes = [U64, PD, PD, i32, PD]
    L.mcgpu_climate5_quant_fixed_seed.argtypes = [U64, PD, PD, i32, PI32]
    L.mcgpu_climate5_i32_fixed_seed.argtypes = [U64, PI32, PI32, i32, PI32]
    L.mcgpu_climate_ctx_create.argtypes = [U64, i32]
    L.mcgpu_climate_ctx_create.restype = VP
    L.mcgpu_climate_ctx_query_i32.argtypes = [VP, PI32, PI32, i32, PI32]
    L.mcgpu_climate_ctx_destroy.argtypes = [VP]
    L.mcgpu_climate_ctx_destroy.restype = None
    L.mcgpu_climate6_i32_fixed_seed.argtypes = [U64, PI32, PI32, PI32, i32, PI32]
    L.mcgpu_climate6_i32_fixed_seed_ex.argtypes = [U64, i32, PI32, PI32, PI32, i32, PI32]
    L.mcgpu_biome20_i32_fixed_seed.argtypes = [U64, PI32, PI32, PI32, i32, PI32]
    L.mcgpu_biome20_ctx_create.argtypes = [U64, i32]
    L.mcgpu_biome20_ctx_create.restype = VP
    L.mcgpu_biome20_ctx_query.argtypes = [VP, PI32, PI32, PI32, i32, PI32]
    L.mcgpu_biome20_ctx_destroy.argtypes = [VP]
    L.mcgpu_biome20_ctx_destroy.restype = None
    for name in ("mcgpu_xoroshiro_nextlong_batch", "mcgpu_xperlin_sample_batch",
                 "mcgpu_xperlin_sample_fixed_seed", "mcgpu_continentalness_noise_fixed_seed",
                 "mcgpu_climate5_fixed_seed", "mcgpu_climate5_quant_fixed_seed",
                 "mcgpu_climate5_i32_fixed_seed", "mcgpu_climate_ctx_query_i32",
                 "mcgpu_climate6_i32_fixed_seed", "mcgpu_climate6_i32_fixed_seed_ex",
                 "mcgpu_biome20_i32_fixed_seed", "mcgpu_biome20_ctx_query"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _rows(buf, total, k):
    n = min(max(total, 0), k)
    return [(buf[i].size, buf[i].center_x, buf[i].center_z, buf[i].obs_count,
             buf[i].afk_x, buf[i].afk_z) for i in range(n)]


def search_slime_clusters_centered(seed, rd_max, min_size, max_size=221, rd_min=0,
                                   search_center_x=0, search_center_z=0, max_results=65536,
                                   threads=0, precise_afk=0):
    """Mirror of SlimeCore.cpp:643-651. Returns (total, rows); total -1 on bad args."""
    buf = (ExtChunkResult * max(1, max_results))()
    total = lib().search_slime_clusters_centered(seed, rd_max, min_size, max_size, rd_min,
                                                 search_center_x, search_center_z, buf,
                                                 max_results, threads, precise_afk)
    return total, _rows(buf, total, max_results)


def search_slime_clusters(seed, rd_max, min_size, max_size=221, rd_min=0, max_results=65536,
                          threads=0, precise_afk=0):
    """Mirror of SlimeCore.cpp:634-641 (search around chunk 0,0)."""
    return search_slime_clusters_centered(seed, rd_max, min_size, max_size, rd_min, 0, 0,
                                          max_results, threads, precise_afk)


def search_slime_clusters_gpu_v1_centered(seed, rd_max, min_size, max_size=221, rd_min=0,
                                          search_center_x=0, search_center_z=0,
                                          max_results=65536, precise_afk=0):
    """Mirror of SlimeCoreGPU.cu:3927-3935."""
    buf = (ExtChunkResult * max(1, max_results))()
    total = lib().search_slime_clusters_gpu_v1_centered(seed, rd_max, min_size, max_size, rd_min,
                                                        search_center_x, search_center_z, buf,
                                                        max_results, precise_afk)
    return total, _rows(buf, total, max_results)


def _row_buffer(rows):
    buf = (ExtChunkResult * max(1, len(rows)))()
    for i, r in enumerate(rows):
        buf[i] = ExtChunkResult(*r)
    return buf


def set_y_scan_config(enabled, platform_y=-64, y_min=-64, y_max=64, y_step=4):
    """Mirror of SlimeCore.cpp:364-371: the AFK heights of the precise refinement."""
    lib().set_y_scan_config(enabled, platform_y, y_min, y_max, y_step)


def score_candidates_precise(seed, rows, scan_xz=1, threads=0):
    """Mirror of SlimeCore.cpp:654-726 on a list of 6-tuples; returns the new rows."""
    buf = _row_buffer(rows)
    lib().score_candidates_precise(seed, buf, len(rows), threads, scan_xz)
    return _rows(buf, len(rows), len(rows))


def refine_candidates_y(seed, rows, threads=0):
    """Mirror of SlimeCoreGPU.cu:3958-3990 on a list of 6-tuples; returns the new rows."""
    buf = _row_buffer(rows)
    lib().refine_candidates_y(seed, buf, len(rows), threads)
    return _rows(buf, len(rows), len(rows))


def refine_multi(seeds, rows_per_world):
    """refine_candidates_y of each world's rows, all worlds in one launch."""
    first = [0]
    for rows in rows_per_world:
        first.append(first[-1] + len(rows))
    flat = [r for rows in rows_per_world for r in rows]
    buf = _row_buffer(flat)
    n = len(seeds)
    rc = lib().refine_candidates_y_multi(n, (C.c_int64 * max(1, n))(*seeds),
                                         (C.c_int32 * (n + 1))(*first), buf)
    if rc != 0:
        raise RuntimeError("refine_candidates_y_multi failed")
    out = _rows(buf, len(flat), len(flat))
    return [out[first[w]:first[w + 1]] for w in range(n)]


def refine_stats():
    """(kernel_ns, evaluations) of the last precise refinement."""
    ns, ev = C.c_int64(), C.c_int64()
    lib().slime_refine_stats(C.byref(ns), C.byref(ev))
    return ns.value, ev.value


def device_info():
    """(compute_units, clock_khz, hbm_bytes) of the current HIP device, or None."""
    cu, clk, mem = C.c_int32(), C.c_int32(), C.c_int64()
    if lib().slime_device_info(C.byref(cu), C.byref(clk), C.byref(mem)) != 0:
        return None
    return cu.value, clk.value, mem.value


KERNEL_PLAIN, KERNEL_BANDED, KERNEL_ALL, KERNEL_BANDED_THREE = 0, 1, 2, 3


VALU_KINDS = ("v_add_u32", "v_xor_b32", "v_mul_lo_u32", "v_bcnt_u32_b32", "v_mad_u32_u24",
              "v_alignbit_b32", "v_xor_b32 literal", "v_add_u32_e64", "v_mul_u32_u24",
              "v_xor_b32 sgpr")


def valu_rate(kind, iters=8192):
    """Issue rate of one kind of integer VALU instruction on this device
    (slime_valu_rate): lane instructions per second, or None on failure."""
    n, ns = C.c_int64(), C.c_int64()
    if lib().slime_valu_rate(kind, iters, C.byref(n), C.byref(ns)) != 0 or ns.value <= 0:
        return None
    return n.value / (ns.value * 1e-9)


WINDOW_VALU, WINDOW_MFMA = 0, 1


def window_ab(method, bands, blocks, steps, min_size, seed=0):
    """Counting-stage A/B (slime_window_ab): (checksum, passes, kernel_ns),
    or None on failure."""
    cs, n, ns = C.c_uint32(), C.c_uint64(), C.c_int64()
    if lib().slime_window_ab(method, bands, blocks, steps, min_size, seed & 0xFFFFFFFF,
                             C.byref(cs), C.byref(n), C.byref(ns)) != 0:
        return None
    return cs.value, n.value, ns.value


def kernel_stats(kind=KERNEL_ALL, reset=False):
    """(launches, kernel_ns, centers) of the scan kernels since the last reset."""
    n, ns, c = C.c_int64(), C.c_int64(), C.c_int64()
    lib().slime_kernel_stats(kind, C.byref(n), C.byref(ns), C.byref(c), 1 if reset else 0)
    return n.value, ns.value, c.value


MODE_AUTO, MODE_PLAIN, MODE_BANDED, MODE_BANDED_SQUARE, MODE_BANDED_UNION = 0, 1, 2, 3, 4
MODE_BANDED_UNION3 = 5  # the union bound with three x-phases past the whole 2^18 blocks


def scan_bands(seed, x0, z0, w, h, min_size, max_size=221, band_rows=0, mode=MODE_AUTO,
               max_results=65536):
    """Scan of [x0,x0+w) x rows of [z0,z0+h) (or the strided bands z0+b*2^17+s,
    s < band_rows) with a forced kernel: (total, rows, hist[222])."""
    buf = (ExtChunkResult * max(1, max_results))()
    hist = (C.c_int64 * 222)()
    total = lib().slime_scan_bands(seed, x0, z0, w, h, band_rows, mode, min_size, max_size,
                                   buf, max_results, hist)
    return total, _rows(buf, total, max_results), list(hist)


def scan_rect(seed, x0, z0, w, h, min_size, max_size=221, max_results=65536):
    """Rectangle scan with the score histogram: (total, rows, hist[222])."""
    buf = (ExtChunkResult * max(1, max_results))()
    hist = (C.c_int64 * 222)()
    total = lib().slime_scan_rect(seed, x0, z0, w, h, min_size, max_size, buf, max_results, hist)
    return total, _rows(buf, total, max_results), list(hist)


def search_batch(seeds, rd_max, min_size, max_size=None, rd_min=None, center_x=None,
                 center_z=None, max_results=1000, use_graph=False, persistent=False):
    """Many searches scanned together (slime_search_batch; persistent: the
    work-queue kernels): a list of (total, rows) per search, each identical to
    search_slime_clusters_gpu_v1_centered."""
    n = len(seeds)
    def arr(t, v, default):
        v = [default] * n if v is None else list(v)
        return (t * max(1, n))(*v)
    a_seed = arr(C.c_int64, seeds, 0)
    a_rd = arr(C.c_int64, rd_max, 0)
    a_min = arr(C.c_int32, min_size, 0)
    a_max = arr(C.c_int32, max_size, 221)
    a_rmin = arr(C.c_int64, rd_min, 0)
    a_cx = arr(C.c_int32, center_x, 0)
    a_cz = arr(C.c_int32, center_z, 0)
    buf = (ExtChunkResult * (max(1, n) * max_results))()
    tot = (C.c_int64 * max(1, n))()
    rc = lib().slime_search_batch(n, a_seed, a_rd, a_min, a_max, a_rmin, a_cx, a_cz, buf,
                                  max_results, tot,
                                  (1 if use_graph else 0) | (2 if persistent else 0))
    if rc != 0:
        raise RuntimeError("slime_search_batch failed")
    out = []
    for i in range(n):
        t = tot[i]
        k = min(max(t, 0), max_results)
        base = i * max_results
        out.append((t, [(buf[base + j].size, buf[base + j].center_x, buf[base + j].center_z,
                         buf[base + j].obs_count, buf[base + j].afk_x, buf[base + j].afk_z)
                        for j in range(k)]))
    return out


def batch_stats():
    """Figures of the last batch as a dict (slime_batch_stats)."""
    v64 = [C.c_int64() for _ in range(5)]
    v32 = [C.c_int32() for _ in range(3)]
    hits = [C.c_int64() for _ in range(2)]
    lib().slime_batch_stats(*[C.byref(x) for x in v64], *[C.byref(x) for x in v32],
                            *[C.byref(x) for x in hits])
    keys64 = ("centers", "kernel_ns", "prepare_ns", "sweep_ns", "wall_ns")
    keys32 = ("launches", "jobs", "fallbacks")
    d = {k: x.value for k, x in zip(keys64, v64)}
    d.update({k: x.value for k, x in zip(keys32, v32)})
    d["hits_stored"], d["hits_copied"] = hits[0].value, hits[1].value
    return d


def sweep_costs():
    """The autotuner's measured table: [(nb, ph, ns_per_tile_row, extra_rows)]."""
    cap = 32
    a, b = (C.c_double * cap)(), (C.c_double * cap)()
    nb, ph = (C.c_int32 * cap)(), (C.c_int32 * cap)()
    n = lib().slime_sweep_costs(a, b, nb, ph, cap)
    if n < 0:
        raise RuntimeError("slime_sweep_costs failed")
    return [(nb[i], ph[i], a[i], b[i]) for i in range(n)]


def batch_class_centers():
    """Centers the last search_batch decided per kernel class, keyed (nb, ph)."""
    cap = 32
    out = (C.c_int64 * cap)()
    n = lib().slime_batch_class_centers(out, cap)
    if n < 0:
        raise RuntimeError("slime_batch_class_centers failed")
    classes = [(c[0], c[1]) for c in sweep_costs()]
    return {classes[i]: out[i] for i in range(min(n, len(classes))) if out[i]}


QUEUED, RUNNING, PAUSED, DONE, CANCELLED, FAILED = range(6)


class Session:
    """Served searches with continuous batching (slime_session_*): requests
    are admitted by device footprint, scanned together in rounds, and can be
    paused, resumed and cancelled while others run. round_ms is the device
    time of one round; rounds are pipelined, so short rounds (the default 10
    ms) cost little throughput."""

    def __init__(self, hbm_budget=4 << 30, max_active=64, round_ms=10, persistent=False):
        self._h = lib().slime_session_create_ex(hbm_budget, max_active, round_ms,
                                                1 if persistent else 0)
        if not self._h:
            raise RuntimeError("slime_session_create failed")

    def close(self):
        if self._h:
            lib().slime_session_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, seed, rd_max, min_size, max_size=221, rd_min=0, center_x=0, center_z=0,
               max_results=1000):
        return lib().slime_session_submit(self._h, seed, rd_max, min_size, max_size, rd_min,
                                          center_x, center_z, max_results)

    def pause(self, rid):
        return lib().slime_session_pause(self._h, rid) == 0

    def resume(self, rid):
        return lib().slime_session_resume(self._h, rid) == 0

    def cancel(self, rid):
        return lib().slime_session_cancel(self._h, rid) == 0

    def poll(self, rid):
        """(state, progress_pct, total, latency_ns)."""
        pct, tot, lat = C.c_int32(), C.c_int64(), C.c_int64()
        st = lib().slime_session_poll(self._h, rid, C.byref(pct), C.byref(tot), C.byref(lat))
        return st, pct.value, tot.value, lat.value

    def wait(self, rid, timeout_ms=-1):
        return lib().slime_session_wait(self._h, rid, timeout_ms)

    def result(self, rid, cap=1000):
        """(total, rows) of a finished request."""
        buf = (ExtChunkResult * max(1, cap))()
        total = lib().slime_session_result(self._h, rid, buf, cap)
        return total, _rows(buf, total, cap)

    def release(self, rid):
        return lib().slime_session_release(self._h, rid) == 0

    def stats(self):
        v64 = [C.c_int64() for _ in range(4)]
        v32 = [C.c_int32() for _ in range(3)]
        lib().slime_session_stats(self._h, C.byref(v64[0]), C.byref(v64[1]), C.byref(v32[0]),
                                  C.byref(v32[1]), C.byref(v32[2]), C.byref(v64[2]),
                                  C.byref(v64[3]))
        return {"centers_done": v64[0].value, "kernel_ns": v64[1].value, "rounds": v32[0].value,
                "active": v32[1].value, "queued": v32[2].value, "hbm_in_use": v64[2].value,
                "hbm_budget": v64[3].value}


DEEP_DARK, MUSHROOM_FIELDS = 1, 2  # filter flags / mask bits
ALL_CHUNKS = 4  # level-range filter: every chunk of the disk, not only its slime chunks


class BiomeContext:
    """One world's overworld biome source on the GPU (mcgpu_modern_ctx_*):
    the candidate filter (Deep Dark / Mushroom Fields under a cluster's slime
    chunks) and biome point queries at quart coordinates."""

    def __init__(self, world_seed, capacity=1 << 16, version=2602, large_biomes=False):
        self._h = lib().mcgpu_modern_ctx_create_ex(version, world_seed & ((1 << 64) - 1),
                                                   1 if large_biomes else 0, capacity)
        if not self._h:
            raise RuntimeError("mcgpu_modern_ctx_create_ex failed")
        self.capacity = capacity

    def close(self):
        if self._h:
            lib().mcgpu_modern_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _arr(v):
        return (C.c_int32 * max(1, len(v)))(*v)

    def filter(self, center_x, center_z, flags=DEEP_DARK | MUSHROOM_FIELDS):
        n = len(center_x)
        out = (C.c_int32 * max(1, n))()
        rc = lib().mcgpu_modern_ctx_filter_candidates(self._h, self._arr(center_x),
                                                      self._arr(center_z), n, flags, out)
        if rc != 0:
            raise RuntimeError(f"filter failed ({rc})")
        return list(out)[:n]

    def filter_levels(self, center_x, center_z, y_min, y_max, y_step,
                      flags=DEEP_DARK | MUSHROOM_FIELDS):
        """The filter over block heights y_min..y_max step y_step (both biomes
        at every height; flags | ALL_CHUNKS samples the whole disk)."""
        n = len(center_x)
        out = (C.c_int32 * max(1, n))()
        rc = lib().mcgpu_modern_ctx_filter_candidates_levels(
            self._h, self._arr(center_x), self._arr(center_z), n, flags, y_min, y_max, y_step, out)
        if rc != 0:
            raise RuntimeError(f"filter_levels failed ({rc})")
        return list(out)[:n]

    def query(self, xs, ys, zs):
        n = len(xs)
        out = (C.c_int32 * max(1, n))()
        rc = lib().mcgpu_modern_ctx_query(self._h, self._arr(xs), self._arr(ys), self._arr(zs), n,
                                          out)
        if rc != 0:
            raise RuntimeError(f"query failed ({rc})")
        return list(out)[:n]

    def columns(self, xs, zs):
        n = len(xs)
        out = (C.c_int32 * max(1, 6 * n))()
        rc = lib().mcgpu_modern_ctx_columns(self._h, self._arr(xs), self._arr(zs), n, out)
        if rc != 0:
            raise RuntimeError(f"columns failed ({rc})")
        return [tuple(out[6 * i:6 * i + 6]) for i in range(n)]

    def last_filter_ms(self):
        return lib().mcgpu_modern_ctx_last_filter_ms(self._h)


def filter_multi(world_seeds, candidates, flags=DEEP_DARK | MUSHROOM_FIELDS, large_biomes=False,
                 version=2602, levels=None):
    """Biome filter of candidates of many worlds in one launch: candidates[w]
    is a list of (center_chunk_x, center_chunk_z) of world world_seeds[w];
    levels = (y_min, y_max, y_step) selects the level-range form. Returns the
    masks per world and the device time in ms."""
    n = len(world_seeds)
    first = [0]
    for c in candidates:
        first.append(first[-1] + len(c))
    flat = [c for cs in candidates for c in cs]
    total = first[-1]
    seeds = (C.c_uint64 * max(1, n))(*[s & ((1 << 64) - 1) for s in world_seeds])
    f = (C.c_int32 * (n + 1))(*first)
    xs = (C.c_int32 * max(1, total))(*[c[0] for c in flat])
    zs = (C.c_int32 * max(1, total))(*[c[1] for c in flat])
    out = (C.c_int32 * max(1, total))()
    large = 1 if large_biomes else 0
    if levels is None:
        rc = lib().mcgpu_filter_multi(n, version, seeds, f, xs, zs, flags, large, out)
    else:
        rc = lib().mcgpu_filter_multi_levels(n, version, seeds, f, xs, zs, flags, large, *levels,
                                             out)
    if rc != 0:
        raise RuntimeError(f"mcgpu_filter_multi failed ({rc})")
    masks = [list(out[first[w]:first[w + 1]]) for w in range(n)]
    return masks, lib().mcgpu_filter_multi_last_ms()


# ---- biome primitives and point queries (mcgpu_* of the reference) -----------------

def _u64(v):
    return (C.c_uint64 * max(1, len(v)))(*[x & ((1 << 64) - 1) for x in v])


def _f64(v):
    return (C.c_double * max(1, len(v)))(*v)


def _i32(v):
    return (C.c_int32 * max(1, len(v)))(*v)


def _check(rc, name):
    if rc != 0:
        raise RuntimeError(f"{name} failed ({rc})")


def xoroshiro_nextlong(seeds, steps):
    """mcgpu_xoroshiro_nextlong_batch: steps outputs per seed."""
    out = (C.c_uint64 * max(1, len(seeds) * steps))()
    _check(lib().mcgpu_xoroshiro_nextlong_batch(_u64(seeds), len(seeds), steps, out), "xoroshiro")
    return [list(out[i * steps:(i + 1) * steps]) for i in range(len(seeds))]


def perlin_sample(seeds, xs, ys, zs):
    """mcgpu_xperlin_sample_batch (one seed per point) or, for an int seed,
    mcgpu_xperlin_sample_fixed_seed."""
    n = len(xs)
    out = (C.c_double * max(1, n))()
    if isinstance(seeds, int):
        rc = lib().mcgpu_xperlin_sample_fixed_seed(seeds & ((1 << 64) - 1), _f64(xs), _f64(ys),
                                                   _f64(zs), n, out)
    else:
        rc = lib().mcgpu_xperlin_sample_batch(_u64(seeds), _f64(xs), _f64(ys), _f64(zs), n, out)
    _check(rc, "xperlin")
    return list(out)[:n]


def continentalness(world_seed, xs, zs):
    n = len(xs)
    out = (C.c_double * max(1, n))()
    _check(lib().mcgpu_continentalness_noise_fixed_seed(world_seed & ((1 << 64) - 1), _f64(xs),
                                                        _f64(zs), n, out), "continentalness")
    return list(out)[:n]


def climate5(world_seed, xs, zs, quantized=False, integer=False):
    """Five climate values (t, h, c, e, w) per point: doubles, quantized at
    double coordinates, or quantized at int32 coordinates (integer=True)."""
    n = len(xs)
    s = world_seed & ((1 << 64) - 1)
    if integer:
        out = (C.c_int32 * max(1, 5 * n))()
        rc = lib().mcgpu_climate5_i32_fixed_seed(s, _i32(xs), _i32(zs), n, out)
    elif quantized:
        out = (C.c_int32 * max(1, 5 * n))()
        rc = lib().mcgpu_climate5_quant_fixed_seed(s, _f64(xs), _f64(zs), n, out)
    else:
        out = (C.c_double * max(1, 5 * n))()
        rc = lib().mcgpu_climate5_fixed_seed(s, _f64(xs), _f64(zs), n, out)
    _check(rc, "climate5")
    return [tuple(out[5 * i:5 * i + 5]) for i in range(n)]


def climate6(world_seed, xs, ys, zs, large_biomes=False):
    """Six quantized values (t, h, c, e, depth, w) per quart point."""
    n = len(xs)
    out = (C.c_int32 * max(1, 6 * n))()
    _check(lib().mcgpu_climate6_i32_fixed_seed_ex(world_seed & ((1 << 64) - 1),
                                                  1 if large_biomes else 0, _i32(xs), _i32(ys),
                                                  _i32(zs), n, out), "climate6")
    return [tuple(out[6 * i:6 * i + 6]) for i in range(n)]


def biome20(world_seed, xs, ys, zs):
    """mcgpu_biome20_i32_fixed_seed: biome ids with the 1.20 tree."""
    n = len(xs)
    out = (C.c_int32 * max(1, n))()
    _check(lib().mcgpu_biome20_i32_fixed_seed(world_seed & ((1 << 64) - 1), _i32(xs), _i32(ys),
                                              _i32(zs), n, out), "biome20")
    return list(out)[:n]
