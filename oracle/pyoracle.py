"""ctypes bindings for the CPU checker (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline legs (the
full-world sample, the reference timing and row check of the radius-2,048
search, the AFK-Y refinement timing) import this module: there it is the
timed CPU baseline and the checker of the GPU's rows, never the GPU path. It loads:
  * liboracle.so - the scalar C restatement in this folder;
  * _ref/libslimecore_ref[_avx2].so - the reference's SlimeCore.cpp compiled
    unmodified by `make -C oracle ref` (present only where it was built).
"""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


class ChunkRow(C.Structure):
    """Result row layout shared by the reference ABI (SlimeCore.cpp:182-189)."""
    _fields_ = [("size", C.c_int32), ("center_x", C.c_int32), ("center_z", C.c_int32),
                ("obs_count", C.c_int32), ("afk_x", C.c_int32), ("afk_z", C.c_int32)]


def build(ref=False, force=False):
    """Compile the C restatement (and the reference, when its tree exists);
    force recompiles what is up to date."""
    targets = ["all"]
    if ref and os.path.isdir(os.environ.get("REF_DIR", "/root/reference")):
        targets.append("ref")
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", HERE] + targets,
                   check=True)


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        i32, i64, u64 = C.c_int32, C.c_int64, C.c_uint64
        lib.oracle_is_slime.argtypes = [i64, i64, i64]
        lib.oracle_is_slime.restype = i32
        lib.oracle_first_draw.argtypes = [i64, i64, i64]
        lib.oracle_first_draw.restype = i32
        lib.oracle_slime_from_state0.argtypes = [u64]
        lib.oracle_slime_from_state0.restype = i32
        lib.oracle_java_next_int.argtypes = [i64]
        lib.oracle_java_next_int.restype = i32
        lib.oracle_x_term.argtypes = [i64]
        lib.oracle_x_term.restype = u64
        lib.oracle_z_term.argtypes = [i64]
        lib.oracle_z_term.restype = u64
        lib.oracle_disk_count.argtypes = [i64, i64, i64]
        lib.oracle_disk_count.restype = i32
        lib.oracle_count_rejected_hits.argtypes = []
        lib.oracle_count_rejected_hits.restype = i64
        lib.oracle_search_hist.argtypes = [i64, i64, i32, i32, i64, i32, i32,
                                           C.POINTER(ChunkRow), i32, C.POINTER(C.c_int64)]
        lib.oracle_search_hist.restype = i64
        lib.oracle_search_rect.argtypes = [i64, i64, i64, i64, i64, i32, i32,
                                           C.POINTER(ChunkRow), i32, C.POINTER(C.c_int64)]
        lib.oracle_search_rect.restype = i64
        lib.oracle_spawnable.argtypes = [i64, i64, i64, i32]
        lib.oracle_spawnable.restype = i32
        lib.oracle_refine_rows.argtypes = [i64, C.POINTER(ChunkRow), i32, i32, i32, i32, i32,
                                           i64, i32, i32]
        lib.oracle_refine_rows.restype = None
        _oracle = lib
    return _oracle


def ref_path(avx2=True):
    name = "libslimecore_ref_avx2.so" if avx2 else "libslimecore_ref.so"
    return os.path.join(HERE, "_ref", name)


def ref_lib(avx2=None):
    """The compiled reference, or None where it was not built."""
    if avx2 is None:
        avx2 = cpu_has_avx2()
    path = ref_path(avx2)
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    i32, i64 = C.c_int32, C.c_int64
    lib.search_slime_clusters_centered.argtypes = [i64, i64, i32, i32, i64, i32, i32,
                                                   C.POINTER(ChunkRow), i32, i32, i32]
    lib.search_slime_clusters_centered.restype = i64
    lib.reset_cancel.argtypes = []
    lib.set_y_scan_config.argtypes = [i32, i32, i32, i32, i32]
    lib.set_y_scan_config.restype = None
    lib.score_candidates_precise.argtypes = [i64, C.POINTER(ChunkRow), i32, i32, i32]
    lib.score_candidates_precise.restype = None
    lib.refine_candidates_y.argtypes = [i64, C.POINTER(ChunkRow), i32, i32]
    lib.refine_candidates_y.restype = None
    return lib


def cpu_has_avx2():
    try:
        with open("/proc/cpuinfo") as f:
            return " avx2" in f.read()
    except OSError:
        return False


def rows_to_tuples(buf, n):
    return [(buf[i].size, buf[i].center_x, buf[i].center_z, buf[i].obs_count,
             buf[i].afk_x, buf[i].afk_z) for i in range(n)]


def oracle_search(seed, rd_max, min_size, max_size=221, rd_min=0, cx=0, cz=0,
                  max_results=65536, hist=False):
    """Returns (total, rows[, hist]) from the C restatement."""
    lib = oracle_lib()
    buf = (ChunkRow * max(1, max_results))()
    h = (C.c_int64 * 222)()
    total = lib.oracle_search_hist(seed, rd_max, min_size, max_size, rd_min, cx, cz,
                                   buf, max_results, h)
    rows = rows_to_tuples(buf, min(max(total, 0), max_results)) if total >= 0 else []
    if hist:
        return total, rows, list(h)
    return total, rows


def ref_search(lib, seed, rd_max, min_size, max_size=221, rd_min=0, cx=0, cz=0,
               max_results=65536, threads=None):
    """Returns (total, rows) from the compiled reference."""
    if threads is None:
        threads = os.cpu_count() or 1
    buf = (ChunkRow * max(1, max_results))()
    lib.reset_cancel()
    total = lib.search_slime_clusters_centered(seed, rd_max, min_size, max_size, rd_min,
                                               cx, cz, buf, max_results, threads, 0)
    rows = rows_to_tuples(buf, min(max(total, 0), max_results)) if total >= 0 else []
    return total, rows


# ---- precise AFK refinement ---------------------------------------------------------

RULE_KEEP_CENTER = 0   # best starts at (0, center, platform y); strictly larger replaces
RULE_FIRST_ORDER = 1   # best starts at -1; ties go to the lowest (offset, level) order


def pack_obs_y(obs, y):
    """obs | (y + 1024) << 20 as a wrapped int32 (SlimeCore.cpp:33-35)."""
    v = (obs | ((y + 1024) << 20)) & 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def y_levels(enabled, platform_y, y_min, y_max, y_step):
    """(y_start, y_step, y_count) of the level list (SlimeCore.cpp:148-159,
    SlimeCoreGPU.cu:78-86); y_step as stored by set_y_scan_config (<= 0 -> 4)."""
    y_step = y_step if y_step > 0 else 4
    start, end, step = (y_min, y_max, y_step) if enabled else (platform_y, platform_y, 1)
    if end < start:
        start, end = end, start
    count = (end - start) // step + 1
    return start, step, max(count, 1)


def oracle_refine(seed, rows, scan_xz, yconf, rule, pack, limit=None):
    """Refines the first `limit` rows (all by default) of a list of 6-tuples."""
    lib = oracle_lib()
    n = len(rows) if limit is None else min(limit, len(rows))
    buf = (ChunkRow * max(1, n))()
    for i in range(n):
        buf[i] = ChunkRow(*rows[i])
    start, step, count = y_levels(*yconf)
    lib.oracle_refine_rows(seed, buf, n, 1 if scan_xz else 0, yconf[1], start, step, count,
                           rule, 1 if pack else 0)
    return rows_to_tuples(buf, n) + list(rows[n:])


def precise_search_rows(family, seed, rows, yconf):
    """Rows of a search called with precise_afk != 0, from its precise_afk == 0
    rows. family "cpu": SlimeCore.cpp:571-629 (100 rows with the y scan, else
    500; every row packed when the y scan is on). family "gpu":
    SlimeCoreGPU.cu:3875-3912 (100 rows; the y scan uses the block kernel's
    rule and packs, otherwise the thread kernel's rule, unpacked)."""
    enabled, platform_y = yconf[0], yconf[1]
    if family == "cpu":
        limit = 100 if enabled else 500
        out = oracle_refine(seed, rows, True, yconf, RULE_KEEP_CENTER, enabled, limit)
        if enabled:
            for i in range(limit, len(out)):
                r = out[i]
                out[i] = r[:3] + (pack_obs_y(r[3], platform_y),) + r[4:]
        return out
    rule = RULE_FIRST_ORDER if enabled else RULE_KEEP_CENTER
    return oracle_refine(seed, rows, True, yconf, rule, enabled, 100)


def ref_refine(lib, fn, seed, rows, yconf, threads=4, scan_xz=1):
    """The compiled reference's score_candidates_precise ("score") or
    refine_candidates_y ("refine_y") over a list of 6-tuples."""
    lib.set_y_scan_config(*yconf)
    lib.reset_cancel()
    buf = (ChunkRow * max(1, len(rows)))()
    for i, r in enumerate(rows):
        buf[i] = ChunkRow(*r)
    if fn == "score":
        lib.score_candidates_precise(seed, buf, len(rows), threads, scan_xz)
    else:
        lib.refine_candidates_y(seed, buf, len(rows), threads)
    return rows_to_tuples(buf, len(rows))


# ---- biome filter checker ----------------------------------------------------------

_biome = None


def biome_lib():
    """libbiome_oracle.so: the C restatement of the biome filter."""
    global _biome
    if _biome is None:
        path = os.path.join(HERE, "libbiome_oracle.so")
        if not os.path.exists(path):
            build()
        _biome = load_biome_lib(path)
    return _biome


def load_biome_lib(path):
    """A build of biome_oracle.c at path, with its argument types set."""
    L = C.CDLL(path)
    u64, i32, i64, P32 = C.c_uint64, C.c_int32, C.c_int64, C.POINTER(C.c_int32)
    L.ob_xoro_stream.argtypes = [u64, C.c_int, C.POINTER(u64)]
    L.ob_xoro_stream.restype = None
    L.ob_perlin_sample_seeded.argtypes = [u64, C.c_double, C.c_double, C.c_double]
    L.ob_perlin_sample_seeded.restype = C.c_double
    L.ob_climate_new.argtypes = [u64, C.c_int]
    L.ob_climate_new.restype = C.c_void_p
    L.ob_climate_free.argtypes = [C.c_void_p]
    L.ob_climate_free.restype = None
    L.ob_column.argtypes = [C.c_void_p, i32, i32, P32]
    L.ob_column.restype = None
    ci = C.c_int
    L.ob_biome_at.argtypes = [C.c_void_p, i32, i32, i32, ci]
    L.ob_biome_at.restype = ci
    L.ob_filter.argtypes = [C.c_void_p, i64, i32, i32, ci, ci]
    L.ob_filter.restype = ci
    L.ob_filter_levels.argtypes = [C.c_void_p, i64, i32, i32, ci, i32, i32, i32, ci]
    L.ob_filter_levels.restype = ci
    L.ob_spline.argtypes = [C.POINTER(C.c_float)]
    L.ob_spline.restype = C.c_float
    L.ob_tree_for_version.argtypes = [ci]
    L.ob_tree_for_version.restype = ci
    L.ob_tree_leaf.argtypes = [P32, ci]
    L.ob_tree_leaf.restype = ci
    L.ob_tree_leaf_from.argtypes = [P32, ci, ci]
    L.ob_tree_leaf_from.restype = ci
    L.ob_tree_first_descent.argtypes = [ci]
    L.ob_tree_first_descent.restype = ci
    L.ob_tree_leaf_exhaustive.argtypes = [P32, ci]
    L.ob_tree_leaf_exhaustive.restype = ci
    L.ob_node_distance.argtypes = [P32, ci, ci]
    L.ob_node_distance.restype = u64
    L.ob_leaf_biome.argtypes = [ci, ci]
    L.ob_leaf_biome.restype = ci
    L.ob_tree_len.argtypes = [ci]
    L.ob_tree_len.restype = ci
    L.ob_tree_family.argtypes = [ci]
    L.ob_tree_family.restype = ci
    L.ob_tree_node.argtypes = [ci, ci]
    L.ob_tree_node.restype = u64
    L.ob_climate5.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.ob_climate5.restype = None
    L.ob_climate6.argtypes = [C.c_void_p, i32, i32, i32, P32]
    L.ob_climate6.restype = None
    L.ob_continentalness.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.ob_continentalness.restype = C.c_double
    return L


# The oracle's trees in order (index = tree argument): file suffix of the
# reference's gpu_noise/tables/btree<suffix>.h and first version code.
BIOME_TREES = [("262", 2602), ("215", 1215), ("21wd", 1214), ("20", 1206), ("19", 1194),
               ("192", 1192), ("18", 1180)]
