"""GPU parity of the banded first-draw scan and its rejection correction.

The banded kernel shares one LCG low half across 2 x-phases (1 for regions
at most 2^17 columns wide) and up to 14 z-bands spaced 2^17 apart. The checks:
  * strided band sets (a few rows of each band) against the CPU oracle run on
    each band's rectangle: totals, whole score histograms and ranked rows,
    for one phase, two phases with a masked edge, and right-aligned tails;
  * the banded scan against the plain per-chunk scan (an independent kernel:
    own tiling, ballot packing and counting) on the same regions:
    bit-identical histograms over up to ~10^11 centers, the size-independent
    check;
  * seeds built so that chunks whose first draw Java rejects lie inside the
    scanned region, both rejected-then-hit and first-draw false positives.
"""
import ctypes as C
import random

import pytest

pytestmark = pytest.mark.gpu

P = 1 << 17
MASK48 = (1 << 48) - 1
JAVA_MUL = 0x5DEECE66D
SCRAMBLE = 0x5DEECE66D ^ 0x3AD8025F
INT32_MAX = (1 << 31) - 1


def s64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def x_term(x):
    u = x & 0xFFFFFFFF
    a = (u * u * 4987142) & 0xFFFFFFFF
    b = (u * 5947611) & 0xFFFFFFFF
    return (a - (1 << 32) if a >> 31 else a) + (b - (1 << 32) if b >> 31 else b)


def z_term(z):
    u = z & 0xFFFFFFFF
    sq = (u * u) & 0xFFFFFFFF
    lin = (u * 389711) & 0xFFFFFFFF
    return (sq - (1 << 32) if sq >> 31 else sq) * 4392871 + (lin - (1 << 32) if lin >> 31 else lin)


def java_verdict(state1):
    """(first-draw verdict, Java verdict) of a first LCG output."""
    bits = state1 >> 17
    fd = bits % 10 == 0
    st, val = state1, bits % 10
    while bits - val + 9 > INT32_MAX:
        st = (st * JAVA_MUL + 11) & MASK48
        bits = st >> 17
        val = bits % 10
    return fd, val == 0


def rejected_state(rng, want_java_hit):
    """A first LCG output in Java's rejection range with the wanted outcome."""
    while True:
        state1 = (((INT32_MAX - 7) << 17) + rng.getrandbits(20)) & MASK48
        fd, java = java_verdict(state1)
        if fd != java and java == want_java_hit:
            return state1


def seed_with_state(state1, x, z, high16):
    inv = pow(JAVA_MUL, -1, 1 << 48)
    s = ((((state1 - 11) * inv) & MASK48) ^ SCRAMBLE)
    return s64(((s - x_term(x) - z_term(z)) & MASK48) | (high16 << 48))


def oracle_bands(oracle, seed, x0, z0, w, h, band_rows, thr, k):
    """Oracle over the strided bands: (total, hist, top-k rows)."""
    lib = oracle.oracle_lib()
    total, hist, rows = 0, [0] * 222, []
    b = 0
    while b * P < h:
        rows_b = min(band_rows, h - b * P)
        buf = (oracle.ChunkRow * k)()
        hh = (C.c_int64 * 222)()
        t = lib.oracle_search_rect(seed, x0, z0 + b * P, w, rows_b, thr, 221, buf, k, hh)
        assert t >= 0
        total += t
        hist = [a + c for a, c in zip(hist, hh)]
        rows += oracle.rows_to_tuples(buf, min(t, k))
        b += 1
    rows.sort(key=lambda r: (-r[0], r[1], r[2]))
    return total, hist, rows[:k]


@pytest.mark.parametrize("case", range(9))
def test_band_sets_vs_oracle(gpu, oracle, case):
    """case 1: a remainder wider than P; case 4: a remainder of 700 columns
    (the right-aligned tail block starts 700 slot columns before its end);
    cases 5, 6: one phase (narrow and exactly P wide); cases 7, 8: one
    two-phase block whose phase-1 columns run past the region's edge."""
    import slime_amd
    rng = random.Random(77 + case)
    seed = s64(rng.getrandbits(64))
    nb = [1, 3, 5, 16, 2, 3, 2, 1, 4][case]
    band_rows = [24, 12, 9, 6, 20, 20, 10, 30, 8][case]
    w = [2 * P, 3 * P + 1000, 2 * P, 4 * P, 2 * P + 700, 600, P, P + 1, P + P // 2 + 3][case]
    x0 = rng.randint(-10 ** 8, 10 ** 8)
    z0 = rng.randint(-10 ** 8, 10 ** 8)
    h = (nb - 1) * P + band_rows
    thr, k = 36, 500
    print("case", case, seed, x0, z0, w, h, band_rows, flush=True)
    total, rows, hist = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows,
                                             slime_amd.MODE_BANDED, k)
    want_total, want_hist, want_rows = oracle_bands(oracle, seed, x0, z0, w, h, band_rows, thr, k)
    assert total == want_total
    assert hist == want_hist
    assert rows == want_rows


@pytest.mark.parametrize("case", range(4))
def test_one_phase_paired_bands_vs_oracle(gpu, oracle, case):
    """One x-phase with its bands paired (bands 2k and 2k + 1 of a column have
    opposite parity: one test decides both), counted by plane words (7 bands,
    union bound; 11 bands, square bound) and by group windows (3 bands), with
    an odd last band."""
    import slime_amd
    rng = random.Random(9090 + case)
    seed = s64(rng.getrandbits(64))
    nb, band_rows, w, mode = [(7, 10, 600, slime_amd.MODE_BANDED),
                              (11, 6, P - 5, slime_amd.MODE_BANDED_SQUARE),
                              (3, 24, 3000, slime_amd.MODE_BANDED),
                              (8, 8, P, slime_amd.MODE_BANDED_UNION)][case]
    x0 = rng.randint(-10 ** 8, 10 ** 8)
    z0 = rng.randint(-10 ** 8, 10 ** 8)
    h = (nb - 1) * P + band_rows
    thr, k = 35, 500
    print("case", case, seed, x0, z0, w, h, band_rows, flush=True)
    total, rows, hist = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows, mode, k)
    want_total, want_hist, want_rows = oracle_bands(oracle, seed, x0, z0, w, h, band_rows, thr, k)
    assert total == want_total
    assert hist == want_hist
    assert rows == want_rows


@pytest.mark.parametrize("case", range(7))
def test_three_phase_band_sets_vs_oracle(gpu, oracle, case):
    """Three x-phases (x, x + P, x + 2P; phase 2 paired across band pairs)
    forced past the last whole 2P block: widths 2P + 700 (three phases over
    700 slot columns, two over the rest of the block), 3P (one 3P block),
    3P + 1000 (a 3P block and a one-phase strip), 5P - 3 and 7P + 5 (two-phase
    blocks before them); odd and even band counts, 6 bands as two groups of
    three; thresholds low enough that the refinement and emits run."""
    import slime_amd
    rng = random.Random(4242 + case)
    seed = s64(rng.getrandbits(64))
    nb = [2, 3, 4, 5, 6, 3, 2][case]
    band_rows = [20, 12, 10, 8, 6, 10, 16][case]
    w = [2 * P + 700, 3 * P, 3 * P + 1000, 5 * P - 3, 2 * P + 5000, 7 * P + 5, 2 * P + P // 2][case]
    x0 = rng.randint(-10 ** 8, 10 ** 8)
    z0 = rng.randint(-10 ** 8, 10 ** 8)
    h = (nb - 1) * P + band_rows
    thr, k = [36, 38, 35, 37, 36, 38, 34][case], 500
    print("case", case, seed, x0, z0, w, h, band_rows, flush=True)
    slime_amd.kernel_stats(slime_amd.KERNEL_BANDED_THREE, reset=True)
    total, rows, hist = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows,
                                             slime_amd.MODE_BANDED_UNION3, k)
    launches, _, centers = slime_amd.kernel_stats(slime_amd.KERNEL_BANDED_THREE)
    assert launches > 0 and centers > 0
    want_total, want_hist, want_rows = oracle_bands(oracle, seed, x0, z0, w, h, band_rows, thr, k)
    assert total == want_total
    assert hist == want_hist
    assert rows == want_rows


@pytest.mark.parametrize("radius,thr", [(140000, 52), (190000, 56), (262143, 58)])
def test_three_phase_equals_plain_large(gpu, radius, thr):
    """Squares between 2P and 4P wide through the forced three-phase layout
    and the plain per-chunk kernel: identical totals, histograms, rows."""
    import slime_amd
    seed = s64(0x0FED_CBA9_8765_4321 * (radius + 3))
    x0 = z0 = -radius
    w = h = 2 * radius + 1
    print("square", radius, thr, flush=True)
    b = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_BANDED_UNION3, 2000)
    p = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_PLAIN, 2000)
    assert b[0] == p[0] and b[2] == p[2] and b[1] == p[1]


@pytest.mark.parametrize("radius,thr", [(3000, 45), (20000, 52), (65536, 54), (70000, 54),
                                        (131100, 55), (200000, 58), (400000, 57)])
def test_banded_equals_plain_large(gpu, radius, thr):
    """Same square through both kernels: identical totals, histograms, rows."""
    import slime_amd
    seed = s64(0x1234_5678_9ABC_DEF0 * (radius + 7))
    x0 = z0 = -radius
    w = h = 2 * radius + 1
    print("square", radius, thr, flush=True)
    b = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_BANDED, 2000)
    p = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_PLAIN, 2000)
    assert b[0] == p[0] and b[2] == p[2] and b[1] == p[1]
    assert b[0] > 0


@pytest.mark.parametrize("java_hit", [True, False])
def test_rejected_chunks_inside_banded_region(gpu, oracle, java_hit):
    """Several Java-rejected first draws planted in scanned bands."""
    import slime_amd
    rng = random.Random(4242 if java_hit else 2424)
    x0, z0, w, band_rows, nb = -5 * P - 333, 7 * P + 11, 2 * P, 40, 3
    h = (nb - 1) * P + band_rows
    for rep in range(3):
        state1 = rejected_state(rng, java_hit)
        b = rng.randrange(nb)
        x = x0 + rng.randrange(w)
        z = z0 + b * P + rng.randrange(band_rows)
        seed = seed_with_state(state1, x, z, rng.getrandbits(16))
        lib = oracle.oracle_lib()
        assert lib.oracle_first_draw(seed, x, z) != lib.oracle_is_slime(seed, x, z)
        thr = max(0, lib.oracle_disk_count(seed, x, z) - 3)
        print("planted", seed, x, z, thr, flush=True)
        total, rows, hist = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows,
                                                 slime_amd.MODE_BANDED, 300)
        want_total, want_hist, want_rows = oracle_bands(oracle, seed, x0, z0, w, h, band_rows,
                                                        thr, 300)
        assert total == want_total
        assert hist == want_hist
        assert rows == want_rows


def test_full_width_search_hits_recount(gpu, oracle):
    """A wide search (main blocks + right-aligned tail block): every reported
    hit recounts exactly on the CPU and the total equals the plain kernel's."""
    import slime_amd
    seed, rd, thr = -7995527694508729151, 300000, 56
    total, rows = slime_amd.search_slime_clusters_centered(seed, rd, thr, 221, 0, 1000, -2000, 100)
    assert total > 0
    lib = oracle.oracle_lib()
    for r in rows:
        cx, cz = (r[1] - 8) // 16, (r[2] - 8) // 16
        assert lib.oracle_disk_count(seed, cx, cz) == r[0]
    p = slime_amd.scan_bands(seed, 1000 - rd, -2000 - rd, 2 * rd + 1, 2 * rd + 1, thr, 221, 0,
                             slime_amd.MODE_PLAIN, 100)
    assert p[0] == total and p[1] == rows


@pytest.mark.parametrize("radius,thr", [(3000, 38), (70000, 50), (140000, 52), (200000, 55),
                                        (150000, 60), (700000, 53)])
def test_union_and_square_bounds_agree(gpu, radius, thr):
    """The rolling union-of-disks bound and the square bound send different
    groups to refinement but must give identical results (and equal the plain
    kernel on the smaller squares)."""
    import slime_amd
    seed = s64(0x0BADC0DE * (radius + thr))
    x0 = z0 = -radius
    w = h = 2 * radius + 1
    print("bounds", radius, thr, flush=True)
    u = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_BANDED_UNION, 3000)
    s = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_BANDED_SQUARE, 3000)
    assert u[0] == s[0] and u[1] == s[1] and u[2] == s[2]
    assert u[0] > 0 or thr >= 56  # high thresholds: hits are rare at these sizes
    if radius <= 70000:
        p = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_PLAIN, 3000)
        assert u[0] == p[0] and u[1] == p[1] and u[2] == p[2]


@pytest.mark.parametrize("radius,corner", [(70000, 0), (70000, 1), (200000, 2), (200000, 3)])
def test_banded_at_int32_edges(gpu, radius, corner):
    """Squares touching the int32 corners of the coordinate space (centers at
    the largest valid offsets, whose disks reach +-(2^31 - 1)): the banded
    kernels, whose phase-1 columns and halo rows run past the edge, against
    the plain kernel; both against the search call."""
    import slime_amd
    seed = s64(0x5EEDC0DE * (radius + corner))
    lo, hi = -(1 << 31) + 8 + radius, INT32_MAX - 8 - radius
    cx = hi if corner & 1 else lo
    cz = hi if corner & 2 else lo
    w = 2 * radius + 1
    print("edge", radius, cx, cz, flush=True)
    b = slime_amd.scan_bands(seed, cx - radius, cz - radius, w, w, 44, 221, 0,
                             slime_amd.MODE_BANDED, 500)
    p = slime_amd.scan_bands(seed, cx - radius, cz - radius, w, w, 44, 221, 0,
                             slime_amd.MODE_PLAIN, 500)
    assert b[0] == p[0] and b[1] == p[1] and b[2] == p[2]
    assert b[0] > 0
    total, rows = slime_amd.search_slime_clusters_gpu_v1_centered(seed, radius, 44, 221, 0, cx,
                                                                  cz, 500)
    assert total == b[0]
    assert [r[0] for r in rows] == [r[0] for r in b[1]]


@pytest.mark.parametrize("nb,w", [(1, 2 * P + 333), (2, 2 * P + 333), (4, 2 * P + 333),
                                  (8, 2 * P + 333), (9, 2 * P + 333), (12, 2 * P + 333),
                                  (13, 2 * P + 333), (3, 30000), (5, 30000)])
def test_union_bound_band_sets_vs_oracle(gpu, oracle, nb, w):
    """Strided band sets through the union-bound kernels against the oracle:
    by group windows (up to 3 band-phases) and by plane words with the corner
    bound (one phase of one band per thread), two x-phases and one; 9, 12 and
    13 bands run as launches of at most 8."""
    import slime_amd
    rng = random.Random(500 + nb + w)
    seed = s64(rng.getrandbits(64))
    band_rows = 16
    x0 = rng.randint(-10 ** 8, 10 ** 8)
    z0 = rng.randint(-10 ** 8, 10 ** 8)
    h = (nb - 1) * P + band_rows
    thr, k = 35, 400
    print("union bands", nb, seed, x0, z0, flush=True)
    total, rows, hist = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows,
                                             slime_amd.MODE_BANDED_UNION, k)
    want_total, want_hist, want_rows = oracle_bands(oracle, seed, x0, z0, w, h, band_rows, thr, k)
    assert total == want_total
    assert hist == want_hist
    assert rows == want_rows
