"""Progress and telemetry of the control ABI after real searches
(SlimeCoreGPU.cu:3063-3116, 3177, 3778-3779, 3840-3842, 3913): progress 1..79
while scanning and 90 at the end, processed centers = the searched square,
scan time > 0, and the introspection of the kernel that ran."""
import pytest

pytestmark = pytest.mark.gpu


def test_counters_after_searches(gpu):
    import slime_amd
    L = gpu
    L.reset_cancel()
    assert L.get_progress() == 0 and L.get_processed_centers() == 0
    total, _ = slime_amd.search_slime_clusters_gpu_v1_centered(42, 1000, 45, 221, 0, 500, -500, 64)
    assert total >= 0
    assert L.get_progress() == 90
    assert L.get_processed_centers() == 2001 * 2001
    assert L.get_gpu_scan_work_ns() > 0
    assert L.get_gpu_v1_algorithm() == 17            # banded, one x-phase (width < 2^17)
    assert L.get_gpu_v1_shape() == 3 and L.get_gpu_v34_shape() == 3
    # counters accumulate until reset_cancel, as in the reference
    slime_amd.search_slime_clusters_centered(42, 70000, 60, 221, 0, 0, 0, 64)
    assert L.get_processed_centers() == 2001 * 2001 + 140001 * 140001
    assert L.get_gpu_v1_algorithm() == 18            # two x-phases
    L.reset_cancel()
    assert L.get_processed_centers() == 0 and L.get_gpu_scan_work_ns() == 0
    assert L.get_progress() == 0


def test_plain_kernel_introspection(gpu):
    import slime_amd
    L = gpu
    L.reset_cancel()
    tot, rows, _ = slime_amd.scan_bands(7, -500, -500, 1001, 1001, 45, 221, 0,
                                        slime_amd.MODE_PLAIN, 16)
    assert tot >= 0
    assert L.get_gpu_v1_algorithm() == 16 and L.get_gpu_v1_shape() == 2
    assert L.get_processed_centers() == 1001 * 1001


def test_refinement_progress(gpu):
    import slime_amd
    L = gpu
    L.reset_cancel()
    _, rows = slime_amd.search_slime_clusters_gpu_v1_centered(42, 600, 44, 221, 0, 0, 0, 8)
    assert rows
    slime_amd.score_candidates_precise(42, rows, 0)
    assert L.get_progress() == 99
    L.reset_cancel()
    slime_amd.refine_candidates_y(42, rows)
    assert L.get_progress() == 99
