"""The reference's own verification cases, run on the GPU through the C ABI.

* cuda_x2_z15/xp_scan.py:554-560, its default exact scan: seed 123456789,
  radius 393,216 around chunk (912,696, -592,827), threshold 60, called as
  xp_scan.reference() calls the GPU library (:537-551: max size 999, rd_min 0,
  65,536 rows, precise AFK on), with its check that get_processed_centers()
  equals (2r+1)^2 (:549). Here also: every hit recounts exactly on the CPU
  oracle; the independent plain per-chunk kernel over the same square gives
  the same total, rows and 222-bin histogram, at threshold 60 and at 50 (about
  25,000 hits, each recounted); the refined rows equal the oracle's
  refinement under the GPU library's rule.
* cuda_x2_z15/x_phase.py:456, the two --verify origins (seed, x0, z0) =
  (123456789, -1875013, 910011) and (-4551278310048043, 2147391234,
  -2147029941): strided band sets of the banded kernel from those origins
  against the oracle. The second lies 92,413 columns from INT32_MAX, so its
  region is one x-phase as wide as int32 allows; the first has two phases
  with a right-aligned tail block.
"""
import ctypes as C

import pytest

from test_gpu_banded import oracle_bands

pytestmark = pytest.mark.gpu

XP_SEED, XP_RADIUS, XP_X, XP_Z, XP_LO = 123456789, 393216, 912696, -592827, 60
CAPACITY = 65536  # xp_scan.py:16


@pytest.fixture(scope="module")
def xp_default(gpu):
    import slime_amd
    L = gpu
    L.reset_cancel()
    print("xp_scan default: banded search", flush=True)
    buf = (slime_amd.ExtChunkResult * CAPACITY)()
    total = L.search_slime_clusters_gpu_v1_centered(XP_SEED, XP_RADIUS, XP_LO, 999, 0, XP_X, XP_Z,
                                                    buf, CAPACITY, 1)
    processed = L.get_processed_centers()
    L.reset_cancel()
    rows = [(r.size, r.center_x, r.center_z, r.obs_count, r.afk_x, r.afk_z)
            for r in buf[:max(0, min(total, CAPACITY))]]
    return total, rows, processed


def test_xp_scan_default_processed_centers(xp_default):
    total, rows, processed = xp_default
    assert 0 <= total <= CAPACITY
    assert processed == (2 * XP_RADIUS + 1) ** 2


def plain_and_banded(slime_amd, thr):
    w = 2 * XP_RADIUS + 1
    x0, z0 = XP_X - XP_RADIUS, XP_Z - XP_RADIUS
    b = slime_amd.scan_bands(XP_SEED, x0, z0, w, w, thr, 999, 0, slime_amd.MODE_BANDED, CAPACITY)
    p = slime_amd.scan_bands(XP_SEED, x0, z0, w, w, thr, 999, 0, slime_amd.MODE_PLAIN, CAPACITY)
    return b, p


@pytest.mark.parametrize("thr", [60, 50])
def test_xp_scan_default_plain_kernel_and_oracle(xp_default, oracle, thr):
    """Both kernels over the square agree on total, rows and histogram; every
    hit (all of them: fewer than 65,536) recounts on the oracle."""
    import slime_amd
    print("xp_scan default: banded and plain scans, thr", thr, flush=True)
    (bt, brows, bhist), (pt, prows, phist) = plain_and_banded(slime_amd, thr)
    assert bt == pt and bhist == phist and brows == prows
    assert sum(bhist) == bt and bt == len(brows)
    lib = oracle.oracle_lib()
    for r in brows:
        cx, cz = (r[1] - 8) // 16, (r[2] - 8) // 16
        assert max(abs(cx - XP_X), abs(cz - XP_Z)) <= XP_RADIUS
        assert lib.oracle_disk_count(XP_SEED, cx, cz) == r[0]
    if thr == XP_LO:
        total, rows, _ = xp_default
        assert total == bt
        assert [r[:3] for r in rows] == [r[:3] for r in brows]
    else:
        assert bt > 10000


def test_xp_scan_default_refined_rows(xp_default, oracle):
    """precise_afk = 1 on the GPU name: the first 100 rows refined by the GPU
    library's rule (SlimeCoreGPU.cu:3875-3912), y scan off."""
    import slime_amd
    total, rows, _ = xp_default
    plain = [r[:3] + (0, r[1], r[2]) for r in rows]
    slime_amd.set_y_scan_config(0, -64, -64, 64, 4)
    assert rows == oracle.precise_search_rows("gpu", XP_SEED, plain, (0, -64, -64, 64, 4))


P = 1 << 17
XPHASE_ORIGINS = [(123456789, -1875013, 910011), (-4551278310048043, 2147391234, -2147029941)]


@pytest.mark.parametrize("case", range(2))
def test_x_phase_verify_origins(gpu, oracle, case):
    """Strided band sets from x_phase.py's --verify origins: totals, whole
    histograms and rows against the oracle per band, and the same region
    through the plain kernel."""
    import slime_amd
    seed, x0, z0 = XPHASE_ORIGINS[case]
    if case == 0:
        w, nb, band_rows = 2 * P + 5000, 3, 16
    else:
        w, nb, band_rows = (1 << 31) - 1 - 8 - x0, 4, 24  # to the int32 edge (halo included)
    h = (nb - 1) * P + band_rows
    thr, k = 36, 2000
    print("x_phase origin", case, seed, x0, z0, w, h, flush=True)
    want = oracle_bands(oracle, seed, x0, z0, w, h, band_rows, thr, k)
    got = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows, slime_amd.MODE_BANDED, k)
    assert got[0] == want[0] and got[2] == want[1] and got[1] == want[2]
    plain = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, band_rows, slime_amd.MODE_PLAIN, k)
    assert plain == got
    assert want[0] > 0
