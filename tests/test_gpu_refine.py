"""Precise AFK refinement on the GPU (refine.hip) against the C restatement:
score_candidates_precise, refine_candidates_y and the searches with
precise_afk != 0, for both selection rules, every y-scan shape and the
row limits of the CPU and GPU libraries. Bit-exact rows."""
import random

import pytest

pytestmark = pytest.mark.gpu

SEED = 0x123456789ABCDEF0
DEFAULT_Y = (0, -64, -64, 64, 4)
Y_CONFIGS = [
    DEFAULT_Y,
    (1, -64, -64, 64, 4),
    (1, -64, -64, 320, 4),
    (1, 10, 100, -50, 7),
    (1, 0, -20, 20, 0),
    (1, 200, -64, 320, 1),
    (1, -64, 100, 300, 10),       # every height out of reach: all counts 0
]


@pytest.fixture
def yscan(gpu):
    import slime_amd
    yield slime_amd.set_y_scan_config
    slime_amd.set_y_scan_config(*DEFAULT_Y)


def case_rows(oracle, n_search=24, n_random=6, seed=5):
    _, rows = oracle.oracle_search(SEED, 600, 40)
    rng = random.Random(seed)
    rows = list(rows[:n_search])
    for _ in range(n_random):
        x = rng.randint(-3_000_000, 3_000_000)
        z = rng.randint(-3_000_000, 3_000_000)
        rows.append((rng.randint(0, 221), x, z, 0, x, z))
    rows.append((50, 2147483640, -2147483640, 0, 2147483640, -2147483640))
    rows.append((50, -2147483648, 2147483647, 0, -2147483648, 2147483647))
    return rows


@pytest.mark.parametrize("yconf", Y_CONFIGS)
@pytest.mark.parametrize("scan_xz", [0, 1])
def test_score_candidates_precise(oracle, yscan, yconf, scan_xz):
    import slime_amd
    rows = case_rows(oracle)
    print("score_candidates_precise", yconf, scan_xz, len(rows), flush=True)
    yscan(*yconf)
    got = slime_amd.score_candidates_precise(SEED, rows, scan_xz)
    want = oracle.oracle_refine(SEED, rows, scan_xz, yconf, oracle.RULE_KEEP_CENTER, yconf[0])
    assert got == want


@pytest.mark.parametrize("yconf", Y_CONFIGS)
def test_refine_candidates_y(oracle, yscan, yconf):
    import slime_amd
    rows = case_rows(oracle, seed=6)
    print("refine_candidates_y", yconf, len(rows), flush=True)
    yscan(*yconf)
    got = slime_amd.refine_candidates_y(SEED, rows)
    want = oracle.oracle_refine(SEED, rows, 1, yconf, oracle.RULE_FIRST_ORDER, True)
    assert got == want


def test_all_zero_rows_follow_each_rule(oracle, yscan):
    """Heights all out of reach: the GPU rule takes point (-16, -16) and the
    first height, the CPU rule keeps the center and the platform height."""
    import slime_amd
    rows = case_rows(oracle, 4, 0)
    yscan(1, -64, 100, 300, 10)
    def wrap32(v):
        return (v + (1 << 31)) % (1 << 32) - (1 << 31)

    for r, s in zip(rows, slime_amd.refine_candidates_y(SEED, rows)):
        assert s == r[:3] + (oracle.pack_obs_y(0, 100), wrap32(r[1] - 16), wrap32(r[2] - 16))
    for r, s in zip(rows, slime_amd.score_candidates_precise(SEED, rows, 1)):
        assert s == r[:3] + (oracle.pack_obs_y(0, -64), r[1], r[2])


@pytest.fixture
def cpu_rule(monkeypatch):
    """refine_candidates_y with SlimeCore's rule (SLIME_REFINE_RULE=cpu)."""
    monkeypatch.setenv("SLIME_REFINE_RULE", "cpu")
    yield


@pytest.mark.parametrize("yconf", [DEFAULT_Y, (1, -64, -64, 320, 4), (1, -64, 100, 300, 10)])
def test_refine_candidates_y_cpu_rule(oracle, yscan, cpu_rule, yconf):
    """Under SLIME_REFINE_RULE=cpu the shared name follows SlimeCore.cpp:728-773:
    the keep-center rule, always packed (all-zero rows keep the center and the
    platform height)."""
    import slime_amd
    rows = case_rows(oracle, seed=7)
    print("refine_candidates_y cpu rule", yconf, len(rows), flush=True)
    yscan(*yconf)
    got = slime_amd.refine_candidates_y(SEED, rows)
    assert got == oracle.oracle_refine(SEED, rows, 1, yconf, oracle.RULE_KEEP_CENTER, True)
    if yconf[2] == 100:  # every height out of reach
        for r, s in zip(rows, got):
            assert s == r[:3] + (oracle.pack_obs_y(0, yconf[1]), r[1], r[2])


def test_refine_candidates_y_cpu_rule_matches_compiled_reference(oracle, yscan, cpu_rule):
    """Where SlimeCore.cpp was built: its refine_candidates_y on the same rows
    (searched rows, random rows and all-zero rows)."""
    import slime_amd
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference SlimeCore.cpp not built")
    for yconf in ((1, -64, -64, 320, 4), (1, -64, 100, 300, 10)):
        rows = case_rows(oracle, seed=8)
        yscan(*yconf)
        got = slime_amd.refine_candidates_y(SEED, rows)
        want = oracle.ref_refine(ref, "refine_y", SEED, rows, yconf)
        ref.set_y_scan_config(*DEFAULT_Y)
        assert got == want


def test_refine_multi_and_noalloc_follow_cpu_rule(oracle, yscan, cpu_rule):
    import slime_amd
    rows = case_rows(oracle, 6, 2, seed=9)
    yscan(1, -64, 100, 300, 10)
    want = oracle.oracle_refine(SEED, rows, 1, (1, -64, 100, 300, 10), oracle.RULE_KEEP_CENTER, True)
    assert slime_amd.refine_multi([SEED], [rows])[0] == want
    buf = slime_amd._row_buffer(rows)
    assert slime_amd.lib().refine_candidates_y_noalloc(SEED, buf, len(rows)) == 0
    assert slime_amd._rows(buf, len(rows), len(rows)) == want


@pytest.mark.parametrize("yconf", [DEFAULT_Y, (1, -64, -64, 320, 4)])
@pytest.mark.parametrize("family", ["cpu", "gpu"])
def test_search_with_precise_afk(oracle, yscan, yconf, family):
    """Row limits (500 / 100 / 100), rules and packing of the two search
    families (SlimeCore.cpp:571-629, SlimeCoreGPU.cu:3875-3912)."""
    import slime_amd
    k = 700
    print("search precise", family, yconf, flush=True)
    yscan(*yconf)
    if family == "cpu":
        total, rows = slime_amd.search_slime_clusters_centered(SEED, 1000, 40, 221, 0, 0, 0, k,
                                                               precise_afk=1)
    else:
        total, rows = slime_amd.search_slime_clusters_gpu_v1_centered(SEED, 1000, 40, 221, 0, 0,
                                                                      0, k, precise_afk=1)
    t2, plain = oracle.oracle_search(SEED, 1000, 40, max_results=k)
    assert total == t2 and len(plain) > 500
    assert rows == oracle.precise_search_rows(family, SEED, plain, yconf)


def test_cpu_names_match_compiled_reference(oracle, yscan):
    """Where the reference's SlimeCore.cpp was built, its precise search rows
    themselves, for the AFK-Y search's height range."""
    import slime_amd
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference SlimeCore.cpp not built")
    yconf = (1, -64, -64, 320, 4)
    k = 300
    yscan(*yconf)
    total, rows = slime_amd.search_slime_clusters_centered(SEED, 1500, 42, 221, 0, 120, -75, k,
                                                           precise_afk=1)
    buf = (oracle.ChunkRow * k)()
    ref.set_y_scan_config(*yconf)
    ref.reset_cancel()
    want_total = ref.search_slime_clusters_centered(SEED, 1500, 42, 221, 0, 120, -75, buf, k, 8, 1)
    ref.set_y_scan_config(*DEFAULT_Y)
    assert total == want_total
    assert rows == oracle.rows_to_tuples(buf, min(total, k))


def test_refine_many_rows(oracle, yscan):
    """A few hundred rows of a wider search at the AFK-Y search's heights."""
    import slime_amd
    _, rows = oracle.oracle_search(SEED, 2500, 44, max_results=160)
    yconf = (1, -64, -64, 320, 4)
    print("refine many", len(rows), flush=True)
    yscan(*yconf)
    got = slime_amd.refine_candidates_y(SEED, rows)
    assert got == oracle.oracle_refine(SEED, rows, 1, yconf, oracle.RULE_FIRST_ORDER, True)
    ns, evals = slime_amd.refine_stats()
    assert evals == len(rows) * 81 * 33 and ns > 0


def test_refine_multi_equals_per_world_calls(oracle, yscan):
    """Rows of several worlds in one launch, each as refine_candidates_y alone;
    empty worlds included."""
    import slime_amd
    rng = random.Random(12)
    seeds, per_world = [], []
    for w in range(6):
        seed = rng.getrandbits(64) - (1 << 63)
        cx, cz = rng.randint(-10**6, 10**6), rng.randint(-10**6, 10**6)
        _, rows = oracle.oracle_search(seed, 400, 38, 221, 0, cx, cz, max_results=20)
        seeds.append(seed)
        per_world.append(rows if w != 2 else [])
    yconf = (1, -64, -64, 320, 4)
    yscan(*yconf)
    got = slime_amd.refine_multi(seeds, per_world)
    for seed, rows, g in zip(seeds, per_world, got):
        assert g == (slime_amd.refine_candidates_y(seed, rows) if rows else [])
        assert g == oracle.oracle_refine(seed, rows, 1, yconf, oracle.RULE_FIRST_ORDER, True)
