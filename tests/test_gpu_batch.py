"""Batched searches (one launch per kernel class over every search's tiles,
or the persistent work-queue kernels) against the
single-search path and the CPU oracle: every search's total and ranked rows
must be identical to its own separate call."""
import random

import pytest

pytestmark = pytest.mark.gpu


def mixed_batch(seed, n):
    """Searches of every schedule class: tiny to two-phase, annuli, ties."""
    rng = random.Random(seed)
    radii = [0, 1, 5, 40, 300, 3000, 20000, 65535, 65536, 70000, 100000, 150000, 140000,
             180000]
    out = []
    for i in range(n):
        rd = radii[i % len(radii)]
        thr = rng.randint(30, 45) if rd < 1000 else rng.randint(48, 56)
        rd_min = rng.choice([0, 0, rd // 3])
        cx = rng.randint(-1_800_000, 1_800_000)
        cz = rng.randint(-1_800_000, 1_800_000)
        out.append((rng.getrandbits(64) - (1 << 63), rd, thr, 221, rd_min, cx, cz))
    return out


def run_batch(searches, k, use_graph=False, persistent=False):
    import slime_amd
    cols = list(zip(*searches))
    return slime_amd.search_batch(cols[0], cols[1], cols[2], cols[3], cols[4], cols[5], cols[6],
                                  max_results=k, use_graph=use_graph, persistent=persistent)


def test_batch_equals_separate_calls(gpu):
    import slime_amd
    searches = mixed_batch(11, 24)
    got = run_batch(searches, 300)
    for s, (total, rows) in zip(searches, got):
        print("search", s, flush=True)
        want = slime_amd.search_slime_clusters_gpu_v1_centered(s[0], s[1], s[2], s[3], s[4],
                                                               s[5], s[6], 300)
        assert (total, rows) == want
    st = slime_amd.batch_stats()
    assert st["fallbacks"] == 0 and st["jobs"] >= len(searches)


def test_batch_persistent_equals_direct(gpu):
    """The work-queue form (persistent kernels claiming tiles) and the
    one-workgroup-per-tile form give the same results, also from a graph."""
    searches = mixed_batch(12, 20)
    direct = run_batch(searches, 200)
    assert run_batch(searches, 200, persistent=True) == direct
    for _ in range(2):
        assert run_batch(searches, 200, use_graph=True, persistent=True) == direct


def test_batch_small_searches_vs_oracle(gpu, oracle):
    """Small searches of the batch against the CPU oracle directly."""
    rng = random.Random(5)
    searches = []
    for i in range(12):
        rd = rng.choice([0, 3, 17, 64, 150])
        searches.append((rng.getrandbits(64) - (1 << 63), rd, rng.randint(28, 40), 221,
                         rng.choice([0, rd // 2]), rng.randint(-10 ** 6, 10 ** 6),
                         rng.randint(-10 ** 6, 10 ** 6)))
    got = run_batch(searches, 500)
    for s, (total, rows) in zip(searches, got):
        want_total, want_rows = oracle.oracle_search(*s, 500)
        assert total == want_total, s
        assert rows == want_rows, s


def test_batch_graph_replay_and_invalid_entries(gpu):
    """Two graph-replayed batches in a row give the plain batch's results;
    invalid entries get -1 without disturbing the others."""
    searches = mixed_batch(23, 12)
    searches[3] = (1, -5, 40, 221, 0, 0, 0)        # negative radius
    searches[7] = (1, 10, 50, 40, 0, 0, 0)         # max_size < min_size
    plain = run_batch(searches, 200)
    assert plain[3][0] == -1 and plain[7][0] == -1
    for _ in range(3):  # the graph is captured once and replayed, with eager batches between
        assert run_batch(searches, 200, use_graph=True) == plain
        assert run_batch(searches, 200) == plain


def test_batch_rejected_chunks(gpu, oracle):
    """Seeds whose Java-rejected first draws fall inside each search."""
    import slime_amd
    from test_gpu_banded import rejected_state, seed_with_state
    rng = random.Random(99)
    lib = oracle.oracle_lib()
    searches = []
    for i in range(6):
        state1 = rejected_state(rng, i % 2 == 0)
        cx, cz = rng.randint(-10 ** 5, 10 ** 5), rng.randint(-10 ** 5, 10 ** 5)
        x, z = cx + rng.randint(-20, 20), cz + rng.randint(-20, 20)
        seed = seed_with_state(state1, x, z, rng.getrandbits(16))
        assert lib.oracle_first_draw(seed, x, z) != lib.oracle_is_slime(seed, x, z)
        thr = max(0, lib.oracle_disk_count(seed, x, z) - 2)
        searches.append((seed, 30, thr, 221, 0, cx, cz))
    got = run_batch(searches, 400)
    for s, (total, rows) in zip(searches, got):
        want_total, want_rows = oracle.oracle_search(*s, 400)
        assert (total, rows) == (want_total, want_rows), s


def test_batch_top_k_selected_on_device(gpu):
    """Dense thresholds and small K: only the hits at or above the score the
    top K reach come back from the device; rows identical to separate calls."""
    import slime_amd
    rng = random.Random(31)
    searches = [(rng.getrandbits(64) - (1 << 63), rd, thr, 221, 0,
                 rng.randint(-10 ** 6, 10 ** 6), rng.randint(-10 ** 6, 10 ** 6))
                for rd, thr in ((3000, 40), (20000, 44), (70000, 48), (150000, 50), (500, 36))]
    for k in (1, 37, 1000):
        got = run_batch(searches, k)
        st = slime_amd.batch_stats()
        print("top-k", k, st["hits_stored"], st["hits_copied"], flush=True)
        assert st["hits_copied"] < st["hits_stored"]
        for s, (total, rows) in zip(searches, got):
            want = slime_amd.search_slime_clusters_gpu_v1_centered(*s, k)
            assert (total, rows) == want, (s, k)


def test_batch_top_k_with_rejected_chunks_vs_oracle(gpu, oracle):
    """A Java-rejected chunk inside each disk of a dense search with K = 5:
    the recounted centers compete with the device-selected hits."""
    from test_gpu_banded import rejected_state, seed_with_state
    rng = random.Random(77)
    lib = oracle.oracle_lib()
    searches = []
    for i in range(4):
        state1 = rejected_state(rng, i % 2 == 1)
        cx, cz = rng.randint(-10 ** 5, 10 ** 5), rng.randint(-10 ** 5, 10 ** 5)
        x, z = cx + rng.randint(-20, 20), cz + rng.randint(-20, 20)
        seed = seed_with_state(state1, x, z, rng.getrandbits(16))
        assert lib.oracle_first_draw(seed, x, z) != lib.oracle_is_slime(seed, x, z)
        searches.append((seed, 300, 34, 221, 0, cx, cz))
    got = run_batch(searches, 5)
    for s, (total, rows) in zip(searches, got):
        assert (total, rows) == oracle.oracle_search(*s, 5), s


def test_batch_selection_larger_than_staging_falls_back(gpu):
    """More hits at or above the K-th score than the 262,144-hit staging
    buffer: every stored hit comes back, rows still identical."""
    import slime_amd
    s = (0x7A11BACC, 100000, 40, 221, 0, 12345, -54321)
    k = 300000
    (total, rows), = run_batch([s], k)
    st = slime_amd.batch_stats()
    assert total > k and st["hits_copied"] == st["hits_stored"]
    assert (total, rows) == slime_amd.search_slime_clusters_gpu_v1_centered(*s, k)


def test_batch_three_phase_regions(gpu):
    """Regions between 2^18 and 3 * 2^17 columns wide: the sweep planner takes
    the three-phase classes (every column decided once, no masked phase), and
    each search equals its separate call."""
    import slime_amd
    rng = random.Random(31)
    searches = [(rng.getrandbits(64) - (1 << 63), rd, rng.randint(52, 58), 221, 0,
                 rng.randint(-1_500_000, 1_500_000), rng.randint(-1_500_000, 1_500_000))
                for rd in (135000, 150000, 170000, 196000)]
    got = run_batch(searches, 100)
    used = slime_amd.batch_class_centers()
    print("class centers", used, flush=True)
    assert sum(v for (nb, ph), v in used.items() if ph == 3) > 0
    for s, (total, rows) in zip(searches, got):
        print("search", s, flush=True)
        want = slime_amd.search_slime_clusters_gpu_v1_centered(s[0], s[1], s[2], s[3], s[4],
                                                               s[5], s[6], 100)
        assert (total, rows) == want
