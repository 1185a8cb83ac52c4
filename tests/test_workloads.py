"""The synthetic workloads of the measured configurations (CPU only): the
splitmix64 stream against its published outputs, and the batched and served
mixes against the distributions they are specified with."""
import math
import os
import subprocess
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
from slime_amd.workloads import (BATCH_STATE, FULL_WORLD_RADIUS, SplitMix64,  # noqa: E402
                                 batched_searches, served_requests)


def test_splitmix64_known_answers():
    # Vigna's splitmix64 (state += 0x9E3779B97F4A7C15, Stafford mix 13):
    # the first output from state 0 is 0xE220A8397B1DCDAF; from state 1 it is
    # the full-world seed 0x910A2DEC89025CC1.
    assert SplitMix64(0).next() == 0xE220A8397B1DCDAF
    assert SplitMix64(1).next() == 0x910A2DEC89025CC1
    g = SplitMix64(1)
    assert g.signed() == 0x910A2DEC89025CC1 - (1 << 64)


def test_batched_searches_follow_their_distribution():
    b = batched_searches(256)
    assert b == batched_searches(256, BATCH_STATE)
    assert len(b) == 256
    # the batch's candidate count is fixed by the start state
    assert sum((2 * s[1] + 1) ** 2 for s in b) == 3_656_978_681_368
    for seed, rd, thr, mx, rmin, cx, cz in b:
        assert 256 <= rd <= 200_000
        assert 50 <= thr <= 62 and mx == 221 and rmin == 0
        assert -1_800_000 <= cx <= 1_800_000 and -1_800_000 <= cz <= 1_800_000
        assert -(1 << 63) <= seed < (1 << 63)
    # log-uniform radius: about half the searches below the geometric mean
    gm = math.sqrt(256 * 200_000)
    below = sum(1 for s in b if s[1] < gm)
    assert 100 < below < 156


def test_served_requests_follow_the_mix():
    reqs = served_requests(2000, 6.0)
    times = [r["arrival"] for r in reqs]
    assert all(b > a for a, b in zip(times, times[1:]))
    assert 2000 / 6.0 * 0.9 < times[-1] < 2000 / 6.0 * 1.1   # Poisson rate 6/s
    small = sum(1 for r in reqs if r["search"][1] <= 4096)
    mid = sum(1 for r in reqs if r["search"][1] == 50_000)
    full = sum(1 for r in reqs if r["search"][1] == FULL_WORLD_RADIUS)
    assert small + mid + full == 2000
    assert 0.66 < small / 2000 < 0.74 and 0.21 < mid / 2000 < 0.29 and 0.03 < full / 2000 < 0.07
    pauses = sum(1 for r in reqs if r["action"] == "pause")
    cancels = sum(1 for r in reqs if r["action"] == "cancel")
    assert 0.08 < pauses / 2000 < 0.12 and 0.03 < cancels / 2000 < 0.07
    for r in reqs:
        assert 50 <= r["search"][2] <= 65
        if r["search"][1] == FULL_WORLD_RADIUS:
            assert r["search"][5:] == (0, 0)
    assert all(r["search"][1] == 200_000 for r in served_requests(2000, 6.0, full_world=False)
               if r["search"][1] > 50_000)


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
