"""The C ABI library on the CPU: it loads, exports every declared symbol, and
its host-side paths (argument validation, control plane, the single-chunk
test) behave like the reference without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT, load_golden

HEADER = os.path.join(ROOT, "include", "slimefinder.h")
BIOME_HEADER = os.path.join(ROOT, "include", "mcgpu.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int64_t|int32_t|int|float|const char\s*\*|void\s*\*|void|bool)"
                       r"\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(native):
    names = declared_functions()
    assert len(names) >= 26
    for name in names:
        assert hasattr(native, name), name


def test_python_mirror_lists_every_export():
    import slime_amd
    assert sorted(slime_amd.EXPORTS) == declared_functions()
    assert sorted(slime_amd.BIOME_EXPORTS) == declared_functions(BIOME_HEADER)


def test_library_exports_every_biome_symbol(native):
    names = declared_functions(BIOME_HEADER)
    assert len(names) == 30
    for name in names:
        assert hasattr(native, name), name
    assert native.mcgpu_api_version() == 1


def test_invalid_arguments_return_minus_one(native):
    import slime_amd
    cases = [c for c in load_golden("small_searches.json") if c["total"] == -1]
    assert len(cases) == 10
    for c in cases:
        k = c["max_results"]
        buf = (slime_amd.ExtChunkResult * max(1, k))()
        got = native.search_slime_clusters_centered(c["seed"], c["rd_max"], c["min_size"],
                                                    c["max_size"], c["rd_min"], c["cx"], c["cz"],
                                                    buf if k > 0 else None, k, 1, 0)
        assert got == -1, c
    buf = (slime_amd.ExtChunkResult * 4)()
    # null buffer, and the GPU names' tighter rd_max limit (SlimeCoreGPU.cu:3169-3175)
    assert native.search_slime_clusters_centered(1, 10, 40, 221, 0, 0, 0, None, 4, 1, 0) == -1
    lim = ((1 << 31) - 1 - 19) // 2
    assert native.search_slime_clusters_gpu_v1_centered(1, lim + 1, 40, 221, 0, 0, 0, buf, 4, 0) == -1
    assert native.search_slime_clusters_gpu_v1(1, -1, 40, 221, 0, buf, 4, 0) == -1
    assert native.search_slime_clusters(1, 10, 50, 40, 0, buf, 4, 1, 0) == -1
    assert native.slime_scan_rect(1, 0, 0, 0, 10, 40, 221, buf, 4, None) == -1


def test_size_limit_returns_minus_one(native):
    """Regions over 2^48 centers (a square of radius above 8,388,607) are
    refused before any device call, by every search entry point."""
    import slime_amd
    buf = (slime_amd.ExtChunkResult * 4)()
    rd = 8_388_608
    assert (2 * rd + 1) ** 2 > 1 << 48 > (2 * rd - 1) ** 2
    assert native.search_slime_clusters_centered(1, rd, 60, 221, 0, 0, 0, buf, 4, 1, 0) == -1
    assert native.search_slime_clusters(1, rd, 60, 221, 0, buf, 4, 1, 0) == -1
    assert native.search_slime_clusters_gpu_v1_centered(1, rd, 60, 221, 0, 0, 0, buf, 4, 0) == -1
    assert native.search_slime_clusters_gpu_v34(1, rd, 60, 221, 0, buf, 4, 0) == -1
    assert native.slime_scan_rect(1, 0, 0, 1 << 24, (1 << 24) + 1, 60, 221, buf, 4, None) == -1
    assert native.slime_scan_bands(1, 0, 0, (1 << 24) + 1, 1 << 24, 8, 0, 60, 221, buf, 4,
                                   None) == -1
    # a batch of only such searches: each total -1, no device work
    assert slime_amd.search_batch([1, 2], [rd, rd + 5], [60, 60], max_results=4) == \
        [(-1, []), (-1, [])]


def test_biome_filter_argument_checks_need_no_device(native):
    """Invalid flags, heights and world lists are refused before any device
    call (MCGPU_ERR_ARG = -1)."""
    import slime_amd
    slime_amd.lib()  # argtypes
    seeds = (C.c_uint64 * 2)(1, 2)
    first = (C.c_int32 * 3)(0, 1, 2)
    a = (C.c_int32 * 2)()
    f = native.mcgpu_filter_multi_levels
    assert f(2, 2602, seeds, first, a, a, 3, 0, -64, 320, 0, a) == -1  # step
    assert f(2, 2602, seeds, first, a, a, 3, 0, 9, 8, 1, a) == -1  # empty range
    assert f(2, 2602, seeds, first, a, a, 3, 0, 0, 4096, 1, a) == -1  # 4097 heights
    assert f(2, 2602, seeds, first, a, a, 4, 0, -64, 0, 4, a) == -1  # no biome
    assert f(2, 2602, seeds, first, a, a, 16 | 1, 0, -64, 0, 4, a) == -1
    assert f(2, 1179, seeds, first, a, a, 3, 0, -64, 0, 4, a) == -1  # no tree before 1.18
    assert f(0, 2602, seeds, first, a, a, 3, 0, -64, 0, 4, a) == -1
    bad_first = (C.c_int32 * 3)(0, 2, 1)
    assert f(2, 2602, seeds, bad_first, a, a, 3, 0, -64, 0, 4, a) == -1
    assert native.mcgpu_filter_multi(2, 2602, seeds, first, a, a, 4, 0, a) == -1
    assert native.mcgpu_modern_ctx_filter_candidates_levels(None, a, a, 1, 3, -64, 0, 4, a) == -1


def test_is_slime_chunk_matches_reference_bits(native):
    for c in load_golden("predicate.json"):
        for i in range(0, 21, 4):
            for j in range(21):
                want = bool((c["rows"][i] >> j) & 1)
                assert native.is_slime_chunk_c(c["seed"], c["cx"] - 10 + i, c["cz"] - 10 + j) == want


def test_control_plane(native):
    native.reset_cancel()
    assert native.get_progress() == 0
    assert native.get_processed_centers() == 0
    native.request_pause()
    native.resume_search()
    native.request_cancel()
    native.reset_cancel()
    native.set_y_scan_config(1, -64, -64, 64, 0)
    native.set_y_scan_config(0, -64, -64, 64, 4)
    # no scan has run in this process: the reference's initial values
    assert native.get_gpu_v1_algorithm() == -1
    assert native.get_gpu_v1_shape() == 0 and native.get_gpu_v34_shape() == 0
    native.set_gpu_v1_rng_override(1)
    assert native.get_gpu_v1_rng() == 1
    native.set_gpu_v1_rng_override(9)
    assert native.get_gpu_v34_rng() == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import slime_amd
    monkeypatch.setattr(slime_amd, "_lib", None)
    monkeypatch.setattr(slime_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError):
        slime_amd.lib()
