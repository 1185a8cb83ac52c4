"""The precise AFK refinement's C restatement (oracle/slime_oracle.c), pinned
against a numpy brute force of the spawnable-block count, the compiled
reference (score_candidates_precise, refine_candidates_y and the search with
precise_afk, SlimeCore.cpp:571-773) and the known answers of the GPU
library's selection rule (SlimeCoreGPU.cu:2799-2861)."""
import random

import numpy as np
import pytest

SEED = 0x123456789ABCDEF0
DEFAULT_Y = (0, -64, -64, 64, 4)
Y_CONFIGS = [
    DEFAULT_Y,                    # platform height only
    (1, -64, -64, 64, 4),         # 33 heights, all within 128 blocks
    (1, -64, -64, 320, 4),        # the AFK-Y search's range: 97 heights, 33 in reach
    (1, 10, 100, -50, 7),         # reversed range
    (1, 0, -20, 20, 0),           # step 0 is stored as 4
    (1, 200, -64, 320, 1),        # 385 heights, 257 in reach
]


def brute_spawnable(oracle, seed, ox, oz, dy):
    """Blocks (x, z) of slime chunks with 576 < d^2 <= 16384, enumerated."""
    lib = oracle.oracle_lib()
    xs = np.arange(ox - 128, ox + 129)
    zs = np.arange(oz - 128, oz + 129)
    cxs = np.floor_divide(xs, 16)
    czs = np.floor_divide(zs, 16)
    ux, uz = np.unique(cxs), np.unique(czs)
    slime = np.array([[lib.oracle_is_slime(seed, int(a), int(b)) for b in uz] for a in ux],
                     dtype=bool)
    cell = slime[np.searchsorted(ux, cxs)][:, np.searchsorted(uz, czs)]
    d2 = (xs[:, None] - ox) ** 2 + (zs[None, :] - oz) ** 2 + dy * dy
    return int((cell & (d2 > 576) & (d2 <= 16384)).sum())


def random_rows(n, seed=5):
    rng = random.Random(seed)
    rows = []
    for _ in range(n):
        x = rng.randint(-3_000_000, 3_000_000)
        z = rng.randint(-3_000_000, 3_000_000)
        rows.append((rng.randint(0, 221), x, z, 0, x, z))
    rows.append((50, 2147483640, -2147483640, 0, 2147483640, -2147483640))
    return rows


def test_spawnable_count_brute_force(oracle):
    rng = random.Random(9)
    points = [(8, 8, 0), (-1000, 533, 17), (5, 5, 128), (5, 5, -128), (7, -9, 24), (7, -9, 23),
              (123457, -99999, -100), (-16, -16, 129)]
    points += [(rng.randint(-10**6, 10**6), rng.randint(-10**6, 10**6), rng.randint(-130, 130))
               for _ in range(6)]
    lib = oracle.oracle_lib()
    for ox, oz, dy in points:
        assert lib.oracle_spawnable(SEED, ox, oz, dy) == brute_spawnable(oracle, SEED, ox, oz, dy)


def ref_or_skip(oracle):
    lib = oracle.ref_lib()
    if lib is None:
        pytest.skip("reference SlimeCore.cpp not built here (make -C oracle ref)")
    return lib


@pytest.fixture
def ref(oracle):
    lib = ref_or_skip(oracle)
    yield lib
    lib.set_y_scan_config(*DEFAULT_Y)


@pytest.mark.parametrize("yconf", Y_CONFIGS)
@pytest.mark.parametrize("scan_xz", [0, 1])
def test_score_candidates_precise_matches_reference(oracle, ref, yconf, scan_xz):
    _, rows = oracle.oracle_search(SEED, 600, 40)
    rows = rows[:24] + random_rows(6)
    want = oracle.ref_refine(ref, "score", SEED, rows, yconf, scan_xz=scan_xz)
    got = oracle.oracle_refine(SEED, rows, scan_xz, yconf, oracle.RULE_KEEP_CENTER, yconf[0])
    assert got == want


@pytest.mark.parametrize("yconf", Y_CONFIGS[:4])
def test_refine_candidates_y_matches_reference(oracle, ref, yconf):
    _, rows = oracle.oracle_search(SEED, 600, 40)
    rows = rows[:16] + random_rows(4, seed=6)
    want = oracle.ref_refine(ref, "refine_y", SEED, rows, yconf)
    got = oracle.oracle_refine(SEED, rows, 1, yconf, oracle.RULE_KEEP_CENTER, True)
    assert got == want


@pytest.mark.parametrize("yconf", [DEFAULT_Y, (1, -64, -64, 320, 4)])
def test_precise_search_rows_match_reference(oracle, ref, yconf):
    """The CPU search with precise_afk: 500 / 100 refined rows, packing of the
    rest with the y scan (SlimeCore.cpp:571-629)."""
    k = 700
    buf = (oracle.ChunkRow * k)()
    ref.set_y_scan_config(*yconf)
    ref.reset_cancel()
    total = ref.search_slime_clusters_centered(SEED, 1000, 40, 221, 0, 0, 0, buf, k, 4, 1)
    want = oracle.rows_to_tuples(buf, min(total, k))
    t2, plain = oracle.oracle_search(SEED, 1000, 40, max_results=k)
    assert t2 == total and len(plain) > 500
    assert oracle.precise_search_rows("cpu", SEED, plain, yconf) == want


def test_first_order_rule_known_answers(oracle):
    """GPU rule: with every height out of reach all counts are 0, and the
    row takes point 0 (-16, -16) and the first height; the CPU rule keeps the
    center and the platform height. Where counts are positive the two rules
    agree (both take the first maximum in (point, height) order)."""
    _, rows = oracle.oracle_search(SEED, 600, 40)
    rows = rows[:6]
    far = (1, -64, 100, 300, 10)          # every height > 128 above the platform
    first = oracle.oracle_refine(SEED, rows, 1, far, oracle.RULE_FIRST_ORDER, True)
    keep = oracle.oracle_refine(SEED, rows, 1, far, oracle.RULE_KEEP_CENTER, True)
    for r, a, b in zip(rows, first, keep):
        assert a == r[:3] + (oracle.pack_obs_y(0, 100), r[1] - 16, r[2] - 16)
        assert b == r[:3] + (oracle.pack_obs_y(0, -64), r[1], r[2])
    near = (1, -64, -64, 64, 4)
    assert (oracle.oracle_refine(SEED, rows, 1, near, oracle.RULE_FIRST_ORDER, True) ==
            oracle.oracle_refine(SEED, rows, 1, near, oracle.RULE_KEEP_CENTER, True))
    for r in oracle.oracle_refine(SEED, rows, 1, near, oracle.RULE_FIRST_ORDER, True):
        assert r[3] & 0xFFFFF > 0


def test_pack_obs_y():
    from oracle import pyoracle
    assert pyoracle.pack_obs_y(0, -1024) == 0
    assert pyoracle.pack_obs_y(5, -64) == 5 | (960 << 20)
    assert pyoracle.pack_obs_y(1, 1024) == 1 - (1 << 31)     # wraps to int32
    assert pyoracle.y_levels(0, 7, -64, 64, 4) == (7, 1, 1)
    assert pyoracle.y_levels(1, 7, 64, -64, 4) == (-64, 4, 33)
    assert pyoracle.y_levels(1, 7, -20, 20, 0) == (-20, 4, 11)
