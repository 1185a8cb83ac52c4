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


def test_served_mix_equals_alone(gpu):
    import slime_amd
    reqs = searches(3, 18, [40, 300, 3000, 20000, 50000, 100000, 150000, 5])
    # a last one large enough (2e12 centers) to be still running when paused
    reqs.append((0x5EEDF00D, 700000, 56, 221, 0, 123456, -654321))
    k = 300
    with slime_amd.Session(hbm_budget=2 << 30, max_active=8, round_ms=5) as s:
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


def test_pause_holds_progress_and_cancel_queued(gpu):
    import slime_amd
    # about 0.2 s of device work, so that the pause lands while it runs
    big = (0x1234567, 400000, 55, 221, 0, 1000, -1000)
    with slime_amd.Session(hbm_budget=1 << 30, max_active=1, round_ms=2) as s:
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
