"""Served searches (continuous batching): requests arriving while others run,
paused, resumed and cancelled, admitted by device footprint. Every finished
result must equal the same search run alone."""
import random
import threading
import time

import pytest

pytestmark = pytest.mark.gpu


def alone(s, k):
    import slime_amd
    return slime_amd.search_slime_clusters_gpu_v1_centered(*s, max_results=k)


def searches(seed, n, radii):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        rd = radii[i % len(radii)]
        thr = rng.randint(30, 42) if rd < 1000 else rng.randint(48, 58)
        out.append((rng.getrandbits(64) - (1 << 63), rd, thr, 221, rng.choice([0, 0, rd // 4]),
                    rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000)))
    return out


@pytest.mark.parametrize("persistent", [False, True])
def test_served_mix_equals_alone(gpu, persistent):
    """Direct rounds (the default) and the persistent work-queue kernels."""
    import slime_amd
    reqs = searches(3, 18, [40, 300, 3000, 20000, 50000, 100000, 150000, 5, 180000])
    # a last one large enough (2e12 centers) to be still running when paused
    reqs.append((0x5EEDF00D, 700000, 56, 221, 0, 123456, -654321))
    k = 300
    with slime_amd.Session(hbm_budget=2 << 30, max_active=8, round_ms=5,
                           persistent=persistent) as s:
        ids = [None] * len(reqs)

        def arrive():
            for i, r in enumerate(reqs):
                ids[i] = s.submit(*r, max_results=k)
                time.sleep(0.004)

        t = threading.Thread(target=arrive)
        t.start()
        t.join()
        assert all(i and i > 0 for i in ids)
        paused = [ids[5], ids[6], ids[-1]]
        # the smaller ones may have finished already
        assert s.pause(ids[-1])
        for rid in paused[:2]:
            assert s.pause(rid) or s.poll(rid)[0] == slime_amd.DONE
        cancelled = ids[13]
        s.cancel(cancelled)
        time.sleep(0.05)
        for rid in paused:
            st = s.poll(rid)[0]
            assert st in (slime_amd.PAUSED, slime_amd.DONE), st
            assert s.resume(rid) or st == slime_amd.DONE
        for i, rid in enumerate(ids):
            st = s.wait(rid, 120000)
            if rid == cancelled:
                assert st in (slime_amd.CANCELLED, slime_amd.DONE)
                continue
            assert st == slime_amd.DONE, (i, st)
            got = s.result(rid, k)
            print("request", i, reqs[i], got[0], flush=True)
            assert got == alone(reqs[i], k), reqs[i]
            state, pct, total, lat = s.poll(rid)
            assert total == got[0] and lat > 0 and pct == 100
        st = s.stats()
        assert st["active"] == 0 and st["queued"] == 0 and st["hbm_in_use"] == 0


def test_admission_by_footprint_queues_and_finishes(gpu):
    """Low thresholds need large hit regions: a small budget admits only a few
    at once; the rest wait in arrival order and still finish exactly."""
    import slime_amd
    rng = random.Random(17)
    reqs = [(rng.getrandbits(64) - (1 << 63), 20000, 40, 221, 0, rng.randint(-10 ** 6, 10 ** 6),
             rng.randint(-10 ** 6, 10 ** 6)) for _ in range(6)]
    with slime_amd.Session(hbm_budget=64 << 20, max_active=64, round_ms=5) as s:
        ids = [s.submit(*r, max_results=200) for r in reqs]
        assert all(i > 0 for i in ids)
        seen_queued = s.stats()["queued"]
        for rid, r in zip(ids, reqs):
            assert s.wait(rid, 120000) == slime_amd.DONE
            assert s.result(rid, 200) == alone(r, 200)
        assert seen_queued > 0
        assert s.stats()["hbm_in_use"] == 0


@pytest.mark.parametrize("persistent", [False, True])
def test_pause_holds_progress_and_cancel_queued(gpu, persistent):
    import slime_amd
    # about 0.2 s of device work, so that the pause lands while it runs
    big = (0x1234567, 400000, 55, 221, 0, 1000, -1000)
    with slime_amd.Session(hbm_budget=1 << 30, max_active=1, round_ms=2,
                           persistent=persistent) as s:
        rid = s.submit(*big, max_results=100)
        queued = s.submit(*big[:1], 100, 50, 221, 0, 0, 0, max_results=10)
        assert s.pause(rid)
        assert s.cancel(queued)
        time.sleep(0.2)
        st1, p1, _, _ = s.poll(rid)
        time.sleep(0.2)
        st2, p2, _, _ = s.poll(rid)
        assert st1 == st2 == slime_amd.PAUSED and p1 == p2
        assert s.wait(queued, 10000) == slime_amd.CANCELLED
        assert s.resume(rid)
        assert s.wait(rid, 120000) == slime_amd.DONE
        assert s.result(rid, 100) == alone(big, 100)
        assert s.release(rid) and s.poll(rid)[0] == -1


def test_submit_rejects_invalid(gpu):
    import slime_amd
    with slime_amd.Session(hbm_budget=256 << 20, max_active=4, round_ms=5) as s:
        assert s.submit(1, -1, 40) == -1
        assert s.submit(1, 10, 50, 40) == -1
        assert s.submit(1, 10, 40, max_results=0) == -1


def test_served_small_requests_equal_the_oracle(gpu, oracle):
    """The served configuration's small requests against the CPU oracle, not
    the same library: radius <= 300 at thresholds 50..65 (and some lower, so
    that results hold hits), a few seeds whose Java-rejected first draws lie
    in the disk of a hit, one at the int32 corner; arriving while a larger
    request runs; some paused and resumed, two cancelled at once."""
    import slime_amd
    from test_gpu_banded import rejected_state, seed_with_state
    rng = random.Random(2026)
    lib = oracle.oracle_lib()
    reqs = []
    for i in range(22):
        rd = rng.choice([0, 5, 40, 120, 300])
        thr = rng.randint(50, 65) if i % 3 else rng.randint(34, 44)
        reqs.append((rng.getrandbits(64) - (1 << 63), rd, thr, 221, rng.choice([0, 0, rd // 3]),
                     rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000)))
    corner = (1 << 31) - 1 - 300 - 8
    reqs.append((0x7FEEDBEEF, 300, 40, 221, 0, corner, -corner))
    for i in range(4):
        state1 = rejected_state(rng, i % 2 == 0)
        cx, cz = rng.randint(-10 ** 6, 10 ** 6), rng.randint(-10 ** 6, 10 ** 6)
        x, z = cx + rng.randint(-20, 20), cz + rng.randint(-20, 20)
        seed = seed_with_state(state1, x, z, rng.getrandbits(16))
        assert lib.oracle_first_draw(seed, x, z) != lib.oracle_is_slime(seed, x, z)
        reqs.append((seed, 40, max(0, lib.oracle_disk_count(seed, x, z) - 2), 221, 0, cx, cz))
    big = (0xB16B00B5, 600_000, 60, 221, 0, -777_777, 555_555)
    k = 500
    with slime_amd.Session(hbm_budget=2 << 30, max_active=16, round_ms=5) as s:
        big_id = s.submit(*big, max_results=k)
        ids = []
        for i, r in enumerate(reqs):
            ids.append(s.submit(*r, max_results=k))
            if i % 7 == 3:
                s.pause(ids[-1])
        cancelled = {ids[1], ids[9]}
        for rid in cancelled:
            s.cancel(rid)
        time.sleep(0.02)
        for i, rid in enumerate(ids):
            if i % 7 == 3:
                assert s.resume(rid) or s.poll(rid)[0] == slime_amd.DONE
        for r, rid in zip(reqs, ids):
            st = s.wait(rid, 120000)
            total, rows = s.result(rid, k)
            want = oracle.oracle_search(*r, k)
            if rid in cancelled and st == slime_amd.CANCELLED:
                # whatever was found: exact counts, at most the full result
                assert total <= want[0] and len(rows) <= len(want[1])
                for row in rows:
                    assert lib.oracle_disk_count(r[0], (row[1] - 8) // 16, (row[2] - 8) // 16) == row[0]
                continue
            assert st == slime_amd.DONE, (r, st)
            print("request", r, total, flush=True)
            assert (total, rows) == want, r
        assert s.wait(big_id, 120000) == slime_amd.DONE
        assert s.result(big_id, k) == alone(big, k)


@pytest.mark.parametrize("persistent", [False, True])
def test_served_cancel_returns_exact_partial_rows(gpu, oracle, persistent):
    """A request cancelled while it runs returns what its finished tiles
    found: every row's count is the exact (Java) count of its disk, the rows
    are ranked, and the total is at most the full search's."""
    import slime_amd
    lib = oracle.oracle_lib()
    req = (0x5EEDCA7, 700_000, 52, 221, 0, 4321, -8765)
    k = 400
    with slime_amd.Session(hbm_budget=2 << 30, max_active=4, round_ms=2,
                           persistent=persistent) as s:
        rid = s.submit(*req, max_results=k)
        deadline = time.time() + 30
        while s.poll(rid)[1] == 0 and time.time() < deadline:
            time.sleep(0.001)
        assert s.cancel(rid)
        st = s.wait(rid, 60000)
        total, rows = s.result(rid, k)
    full_total, full_rows = alone(req, k)
    print("cancelled", st, total, len(rows), "full", full_total, flush=True)
    assert st == slime_amd.CANCELLED
    assert 0 <= total <= full_total and len(rows) <= min(total, k)
    assert rows == sorted(rows, key=lambda r: (-r[0], r[1], r[2]))
    for row in rows:
        assert lib.oracle_disk_count(req[0], (row[1] - 8) // 16, (row[2] - 8) // 16) == row[0]


FULL_WORLD = (0x910A2DEC89025CC1 - (1 << 64), 1_875_000, 60, 221, 0, 0, 0)


def test_served_full_world_paused_equals_alone(gpu):
    """The served mix's largest request, the full world (14 T centers), in a
    session: paused while it runs (its progress holds while smaller requests
    at thresholds up to 65 arrive and finish), resumed, and equal to the same
    search run alone."""
    import slime_amd
    k = 1000
    small = [(0xABCDEF + i, rd, thr, 221, 0, 1000 * i, -2000 * i)
             for i, (rd, thr) in enumerate([(50_000, 65), (4096, 64), (256, 50), (50_000, 58)])]
    with slime_amd.Session(hbm_budget=8 << 30, max_active=8, round_ms=20) as s:
        rid = s.submit(*FULL_WORLD, max_results=k)
        deadline = time.time() + 60
        while s.poll(rid)[1] == 0 and time.time() < deadline:
            time.sleep(0.002)
        assert s.pause(rid)
        time.sleep(0.1)
        st1, p1, _, _ = s.poll(rid)
        ids = [s.submit(*r, max_results=k) for r in small]
        for r, i in zip(small, ids):
            assert s.wait(i, 60000) == slime_amd.DONE
            assert s.result(i, k) == alone(r, k), r
        st2, p2, _, _ = s.poll(rid)
        print("full world paused at", p1, p2, flush=True)
        assert st1 == st2 == slime_amd.PAUSED and p1 == p2 and 0 < p1 < 100
        assert s.resume(rid)
        assert s.wait(rid, 120000) == slime_amd.DONE
        got = s.result(rid, k)
    assert got[0] == 12
    assert got == alone(FULL_WORLD, k)


def test_pipelined_short_rounds(gpu, oracle):
    """1 ms rounds, so that the next round is nearly always queued while the
    current one runs: requests admitted between rounds, one paused and
    resumed, one cancelled while its tiles are in flight. Finished results
    equal the same search alone; the cancelled one holds exact rows of the
    tiles that finished."""
    import slime_amd
    lib = oracle.oracle_lib()
    reqs = searches(0x51DE, 12, [300, 3000, 40_000, 150_000])
    k = 300
    with slime_amd.Session(hbm_budget=2 << 30, max_active=8, round_ms=1) as s:
        ids = []
        for i, r in enumerate(reqs):
            ids.append(s.submit(*r, max_results=k))
            time.sleep(0.002)
            if i == 5:
                s.pause(ids[2])
        victim = ids[11]
        deadline = time.time() + 30
        while s.poll(victim)[1] == 0 and time.time() < deadline:
            time.sleep(0.0005)
        assert s.cancel(victim)
        assert s.resume(ids[2]) or s.poll(ids[2])[0] == slime_amd.DONE
        for r, rid in zip(reqs, ids):
            st = s.wait(rid, 120000)
            total, rows = s.result(rid, k)
            if rid == victim and st == slime_amd.CANCELLED:
                full = alone(r, k)
                print("cancelled at", total, "of", full[0], flush=True)
                assert total <= full[0] and len(rows) <= min(total, k)
                for row in rows:
                    assert lib.oracle_disk_count(r[0], (row[1] - 8) // 16,
                                                 (row[2] - 8) // 16) == row[0]
                continue
            assert st == slime_amd.DONE, (r, st)
            assert (total, rows) == alone(r, k), r
        rounds = s.stats()["rounds"]
        print("rounds", rounds, flush=True)
        # a round is at least one wave of each class's tiles (a few ms here)
        assert rounds > 8


def test_session_closed_while_rounds_in_flight(gpu):
    """Closing a session while a large request runs: the worker lets the
    rounds in flight end and stops (no hang, memory released), and a new
    session on the same device serves requests as before."""
    import slime_amd
    big = (0x5E55C105E, 1_200_000, 60, 221, 0, 12345, -54321)
    small = (0xC105ED, 3000, 50, 221, 0, -999, 777)
    t0 = time.time()
    with slime_amd.Session(hbm_budget=4 << 30, max_active=4, round_ms=10) as s:
        rid = s.submit(*big, max_results=100)
        deadline = time.time() + 30
        while s.poll(rid)[1] == 0 and time.time() < deadline:
            time.sleep(0.002)
        assert s.poll(rid)[0] == slime_amd.RUNNING
    closed = time.time() - t0
    print("closed after", closed, "s", flush=True)
    assert closed < 10
    with slime_amd.Session(hbm_budget=1 << 30, max_active=4, round_ms=10) as s:
        rid = s.submit(*small, max_results=100)
        assert s.wait(rid, 60000) == slime_amd.DONE
        assert s.result(rid, 100) == alone(small, 100)
