"""refine_candidates_y_noalloc, the entry point that allocates nothing, called
through the C ABI: two calls in a row (the second with fewer rows, so stale
rows of the first stay in the static device array) each equal
refine_candidates_y on the same rows, bit for bit; counts over the array's
65,536 rows are refused with -2, a null buffer with -1."""
import ctypes as C
import random

import pytest

pytestmark = pytest.mark.gpu

AFK_Y = (1, -64, -64, 320, 4)


def candidate_rows(slime_amd, seed, n):
    total, rows = slime_amd.search_slime_clusters_gpu_v1_centered(seed, 3000, 38, 221, 0,
                                                                  1234, -4321, max_results=n)
    assert total >= n
    return rows[:n]


def as_buffer(slime_amd, rows):
    buf = (slime_amd.ExtChunkResult * max(1, len(rows)))()
    for i, r in enumerate(rows):
        buf[i] = slime_amd.ExtChunkResult(*r)
    return buf


def rows_of(buf, n):
    return [(b.size, b.center_x, b.center_z, b.obs_count, b.afk_x, b.afk_z) for b in buf[:n]]


def test_noalloc_equals_refine_candidates_y(gpu):
    import slime_amd
    L = gpu
    slime_amd.set_y_scan_config(*AFK_Y)
    try:
        for seed, n in ((0x5EEDA11C, 3000), (-0x77777777, 700)):
            rows = candidate_rows(slime_amd, seed, n)
            rng = random.Random(n)
            # a few rows moved to other places and the int32 edge
            for i in range(0, n, 97):
                cx = rng.choice([rng.randint(-10 ** 6, 10 ** 6), 134217720, -134217720])
                rows[i] = (rows[i][0], cx * 16 + 8, rows[i][2], 0, cx * 16 + 8, rows[i][2])
            want = as_buffer(slime_amd, rows)
            L.refine_candidates_y(seed, want, n, 0)
            got = as_buffer(slime_amd, rows)
            print("noalloc call", seed, n, flush=True)
            assert L.refine_candidates_y_noalloc(seed, got, n) == 0
            assert rows_of(got, n) == rows_of(want, n)
            assert rows_of(got, n) != rows  # the refinement changed them
    finally:
        slime_amd.set_y_scan_config(0, -64, -64, 64, 4)


def test_noalloc_limits(gpu):
    import slime_amd
    L = gpu
    buf = (slime_amd.ExtChunkResult * 1)()
    assert L.refine_candidates_y_noalloc(1, buf, 65537) == -2
    assert L.refine_candidates_y_noalloc(1, None, 5) == -1
    assert L.refine_candidates_y_noalloc(1, buf, -1) == -1
    assert L.refine_candidates_y_noalloc(1, buf, 0) == 0
