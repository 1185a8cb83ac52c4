"""Pins the CPU oracle (CPU only).

Checks, each chosen so a plausible slip (wrong constant, sign, index or a
dropped term) fails at least one:
  * Java known answers: new Random(0/42).nextInt() (published values);
  * per-chunk bits against the reference's own build_chunk_cache output
    (tests/golden/predicate.json), incl. constructed rejection-path states;
  * the rejected-then-hit state count 104,848 fixed by the reference's rare
    table generator (cuda_x2_z15/generate_slime_rare_tables.py:28);
  * the disk: 221 cells and 68 corner cells of the 17x17 square
    (SlimeCore.cpp:521-523);
  * searches against the reference's search_slime_clusters_centered output
    (tests/golden/c1_hits.json, small_searches.json);
  * brute-force disk counts on tiny regions against the search's ranking.
"""
import random

from conftest import load_golden

MASK48 = (1 << 48) - 1


def test_java_known_answers(oracle):
    lib = oracle.oracle_lib()
    assert lib.oracle_java_next_int(0) == -1155484576
    assert lib.oracle_java_next_int(42) == -1170105035


def test_predicate_matches_reference_bits(oracle):
    lib = oracle.oracle_lib()
    cases = load_golden("predicate.json")
    assert len(cases) > 100
    for c in cases:
        for i in range(21):
            for j in range(21):
                want = (c["rows"][i] >> j) & 1
                got = lib.oracle_is_slime(c["seed"], c["cx"] - 10 + i, c["cz"] - 10 + j)
                assert got == want, (c["kind"], c["seed"], c["cx"] - 10 + i, c["cz"] - 10 + j)


def test_constructed_rejections_reach_redraw(oracle):
    """The constructed cases really put next(31) in the rejected range."""
    lib = oracle.oracle_lib()
    cases = [c for c in load_golden("predicate.json") if c["kind"].startswith("reject")]
    assert len(cases) == 24
    differing = 0
    for c in cases:
        j = int(c["kind"][len("reject"):])
        fd = lib.oracle_first_draw(c["seed"], c["cx"], c["cz"])
        # next(31) = 2^31-8+j; only j == 0 is a multiple of 10 on the first draw
        assert fd == (1 if j == 0 else 0)
        differing += fd != lib.oracle_is_slime(c["seed"], c["cx"], c["cz"])
    assert differing > 0


def test_rejected_hit_count(oracle):
    assert oracle.oracle_lib().oracle_count_rejected_hits() == 104848


def test_disk_shape():
    cells = [(dx, dz) for dz in range(-8, 9) for dx in range(-8, 9) if dx * dx + dz * dz <= 68]
    assert len(cells) == 221
    assert 17 * 17 - len(cells) == 68
    widths = [max(dx for dx, z in cells if z == dz) for dz in range(-8, 9)]
    assert widths == [2, 4, 5, 6, 7, 7, 8, 8, 8, 8, 8, 7, 7, 6, 5, 4, 2]


def test_c1_matches_reference(oracle):
    g = load_golden("c1_hits.json")
    total, rows = oracle.oracle_search(g["seed"], g["rd_max"], g["min_size"], g["max_size"],
                                       g["rd_min"], g["cx"], g["cz"], g["max_results"])
    assert total == g["total"] == 2676
    assert [list(r[:3]) for r in rows] == g["rows"]
    assert all(r[3] == 0 and r[4] == r[1] and r[5] == r[2] for r in rows)


def test_small_searches_match_reference(oracle):
    cases = load_golden("small_searches.json")
    for c in cases:
        total, rows = oracle.oracle_search(c["seed"], c["rd_max"], c["min_size"], c["max_size"],
                                           c["rd_min"], c["cx"], c["cz"], c["max_results"])
        assert total == c["total"], c
        assert [list(r[:3]) for r in rows] == c["rows"], c


def test_search_counts_against_brute_force(oracle):
    """Ranked sizes equal independent per-center disk counts (tiny region)."""
    lib = oracle.oracle_lib()
    rng = random.Random(7)
    for _ in range(3):
        seed = rng.getrandbits(63)
        cx, cz = rng.randint(-10 ** 6, 10 ** 6), rng.randint(-10 ** 6, 10 ** 6)
        total, rows = oracle.oracle_search(seed, 6, 0, 221, 0, cx, cz, 1000)
        assert total == 13 * 13
        brute = sorted(((lib.oracle_disk_count(seed, x, z), x, z)
                        for x in range(cx - 6, cx + 7) for z in range(cz - 6, cz + 7)),
                       key=lambda t: (-t[0], t[1], t[2]))
        assert [(r[0], (r[1] - 8) // 16, (r[2] - 8) // 16) for r in rows] == brute
