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
    # 2^18 - 1 columns: all but one slot column hold both phases
    slime_amd.search_slime_clusters_centered(42, 131071, 60, 221, 0, 0, 0, 64)
    assert L.get_processed_centers() == 2001 * 2001 + 262143 * 262143
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


def test_pause_and_cancel_a_running_search(gpu, oracle):
    """request_pause holds a running search between launches (the processed
    centers stop moving), request_cancel ends it early with the hits found so
    far ranked, each with its exact (Java) count, and reset_cancel lets the
    next search run in full (SlimeCore.cpp:341-362, SlimeCoreGPU.cu:3085-3108,
    3754-3757)."""
    import threading
    import time
    import slime_amd
    L = gpu
    L.reset_cancel()
    rd, thr = 1_500_000, 58        # about a second of scanning in launches of ~70 ms
    full = (2 * rd + 1) ** 2
    out = {}

    def run():
        out["r"] = slime_amd.search_slime_clusters_gpu_v1_centered(0x5EED, rd, thr, 221, 0, 0, 0,
                                                                   max_results=1000)

    t = threading.Thread(target=run)
    t.start()
    try:
        deadline = time.time() + 60
        while L.get_processed_centers() == 0 and t.is_alive() and time.time() < deadline:
            time.sleep(0.002)
        assert t.is_alive(), "the search ended before it could be paused"
        L.request_pause()
        time.sleep(0.3)            # the launch in flight completes
        p1 = L.get_processed_centers()
        time.sleep(0.3)
        p2 = L.get_processed_centers()
        assert p1 == p2 and t.is_alive()
        L.request_cancel()
        t.join(60)
        assert not t.is_alive()
    finally:
        L.request_cancel()
        t.join(120)
        done = L.get_processed_centers()
        L.reset_cancel()
    total, rows = out["r"]
    assert 0 < done < full
    assert total >= len(rows)
    assert rows == sorted(rows, key=lambda r: (-r[0], r[1], r[2]))
    assert all(thr <= r[0] <= 221 for r in rows)
    lib = oracle.oracle_lib()
    for r in rows:
        assert lib.oracle_disk_count(0x5EED, (r[1] - 8) // 16, (r[2] - 8) // 16) == r[0]
    assert L.get_processed_centers() == 0
    total2, _ = slime_amd.search_slime_clusters_gpu_v1_centered(7, 3000, 45, 221, 0, 0, 0, 100)
    assert total2 >= 0 and L.get_processed_centers() == 6001 * 6001 and L.get_progress() == 90
