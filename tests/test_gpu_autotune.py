"""The per-search engine's measured band grouping (searches of at least 2^37
centers; SLIME_BAND_TUNE=0 keeps the even grouping): identical totals,
222-bin histograms and rows either way, every center decided once, on
regions whose last band is partial (the plan may scan it as a short group of
its own) and on a region of whole bands."""
import pytest

pytestmark = pytest.mark.gpu

P = 1 << 17
CASES = [
    # (seed, x0, z0, w, h, min_size)
    (0x51A11, -200_000, -200_000, 400_001, 400_001, 50),          # 3 bands + 6,785 rows
    (-987654321, 1_000_000, -3_000_000, 2 * P * 3 + 777, 9 * P + 50_000, 58),  # 9 bands + a part
    (42, -P, -4 * P, 4 * P, 8 * P, 56),                           # whole bands, 8 of them
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_tuned_plan_equals_even_grouping(gpu, monkeypatch, case):
    import slime_amd
    seed, x0, z0, w, h, thr = CASES[case]
    assert w * h >= 1 << 37
    print("autotune case", case, seed, x0, z0, w, h, thr, flush=True)
    L = slime_amd.lib()
    monkeypatch.setenv("SLIME_BAND_TUNE", "1")
    L.reset_cancel()  # the processed-centers counter runs until reset, as in the reference
    tuned = slime_amd.scan_rect(seed, x0, z0, w, h, thr, 221, 5000)
    assert L.get_processed_centers() == w * h
    monkeypatch.setenv("SLIME_BAND_TUNE", "0")
    L.reset_cancel()
    even = slime_amd.scan_rect(seed, x0, z0, w, h, thr, 221, 5000)
    assert L.get_processed_centers() == w * h
    assert tuned[0] == even[0] and tuned[2] == even[2] and tuned[1] == even[1]
    assert sum(tuned[2]) == tuned[0] and tuned[0] > 0


def test_tuned_plan_takes_three_phases(gpu, monkeypatch):
    """A region 2P + 117,857 columns wide: the measured costs take the
    three-phase layout past the first 2P block (no masked columns), with the
    same totals, histogram and rows as the even two-phase plan."""
    import slime_amd
    seed, x0, z0, w, h, thr = 0x3F3F3F, -190_000, -190_000, 380_001, 380_001, 55
    assert w * h >= 1 << 37
    monkeypatch.setenv("SLIME_BAND_TUNE", "1")
    slime_amd.kernel_stats(slime_amd.KERNEL_BANDED_THREE, reset=True)
    tuned = slime_amd.scan_rect(seed, x0, z0, w, h, thr, 221, 5000)
    launches, _, centers = slime_amd.kernel_stats(slime_amd.KERNEL_BANDED_THREE)
    print("three-phase launches", launches, "centers", centers, flush=True)
    assert launches > 0 and centers > 0
    monkeypatch.setenv("SLIME_BAND_TUNE", "0")
    even = slime_amd.scan_rect(seed, x0, z0, w, h, thr, 221, 5000)
    assert tuned[0] == even[0] and tuned[2] == even[2] and tuned[1] == even[1]
    assert sum(tuned[2]) == tuned[0] and tuned[0] > 0


def test_tuned_plan_pairs_one_phase_columns(gpu, monkeypatch):
    """A region 1.6 P wide: where the phase-1 columns would fall outside, the
    measured plan may scan one phase with paired bands instead of a masked
    two-phase block; the same totals, histogram and rows either way."""
    import slime_amd
    seed, w, h, thr = -0x77AA55, int(1.6 * P) | 1, 2 * P - 3000, 54
    x0 = z0 = -(w // 2)
    monkeypatch.setenv("SLIME_BAND_TUNE", "1")
    tuned = slime_amd.scan_rect(seed, x0, z0, w, h, thr, 221, 5000)
    monkeypatch.setenv("SLIME_BAND_TUNE", "0")
    even = slime_amd.scan_rect(seed, x0, z0, w, h, thr, 221, 5000)
    plain = slime_amd.scan_bands(seed, x0, z0, w, h, thr, 221, 0, slime_amd.MODE_PLAIN, 5000)
    assert tuned[0] == even[0] == plain[0]
    assert tuned[2] == even[2] == plain[2] and tuned[1] == even[1] == plain[1]
