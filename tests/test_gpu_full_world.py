"""Full-world parity at the benchmark's configuration (seed splitmix64(1) =
0x910A2DEC89025CC1, radius 1,875,000 chunks around (0,0), threshold 60:
14,062,507,500,001 centers), where a CPU run of the whole is out of reach:
  1. every reported hit recounts exactly on the CPU oracle;
  2. an independent GPU pass (the plain per-chunk kernel: one LCG evaluation
     per chunk, no phase or band pairing) gives the same total, ranked rows
     and 222-bin score histogram over all 14 T centers;
  3. the CPU oracle over windows around every hit and over random windows
     gives the GPU's counts there."""
import ctypes as C
import random

import pytest

pytestmark = pytest.mark.gpu

SEED = 0x910A2DEC89025CC1 - (1 << 64)
RADIUS = 1_875_000
THR = 60


@pytest.fixture(scope="module")
def full_world(gpu):
    import slime_amd
    print("full world: banded search", flush=True)
    return slime_amd.search_slime_clusters_centered(SEED, RADIUS, THR, 221, 0, 0, 0, 1000)


def test_hits_recount_on_the_oracle(full_world, oracle):
    total, rows = full_world
    assert total == len(rows) == 12
    lib = oracle.oracle_lib()
    for r in rows:
        cx, cz = (r[1] - 8) // 16, (r[2] - 8) // 16
        assert abs(cx) <= RADIUS and abs(cz) <= RADIUS
        assert lib.oracle_disk_count(SEED, cx, cz) == r[0]
    assert rows == sorted(rows, key=lambda r: (-r[0], r[1], r[2]))


def test_independent_plain_kernel_pass(full_world):
    import slime_amd
    total, rows = full_world
    w = 2 * RADIUS + 1
    print("full world: banded and plain scans", flush=True)
    b = slime_amd.scan_bands(SEED, -RADIUS, -RADIUS, w, w, THR, 221, 0, slime_amd.MODE_BANDED,
                             1000)
    p = slime_amd.scan_bands(SEED, -RADIUS, -RADIUS, w, w, THR, 221, 0, slime_amd.MODE_PLAIN,
                             1000)
    assert p[0] == b[0] == total
    assert p[2] == b[2]
    assert sum(p[2]) == total
    assert p[1] == b[1] == rows


def test_oracle_windows(full_world, oracle):
    """64 x 64 centers around each hit and eight random 256 x 256 windows of
    the world: totals, histograms and rows of the GPU's rectangle scan equal
    the oracle's (threshold 50 inside the windows, so they hold hits)."""
    import slime_amd
    _, rows = full_world
    rng = random.Random(60)
    windows = [((r[1] - 8) // 16 - 32, (r[2] - 8) // 16 - 32, 64, 64) for r in rows]
    for _ in range(8):
        windows.append((rng.randint(-RADIUS, RADIUS - 256), rng.randint(-RADIUS, RADIUS - 256),
                        256, 256))
    lib = oracle.oracle_lib()
    k = 4096
    for x0, z0, w, h in windows:
        total, got_rows, hist = slime_amd.scan_rect(SEED, x0, z0, w, h, 50, 221, k)
        obuf = (oracle.ChunkRow * k)()
        ohist = (C.c_int64 * 222)()
        want = lib.oracle_search_rect(SEED, x0, z0, w, h, 50, 221, obuf, k, ohist)
        assert total == want, (x0, z0)
        assert hist == list(ohist)
        assert got_rows == oracle.rows_to_tuples(obuf, min(want, k))
