"""The biome filter's CPU checker (oracle/biome_oracle.c) and the data it
shares with the kernels (slime_amd/csrc/biome_data.inc), pinned without the
reference's CUDA build:
  * noise seeds: MD5 digests of Minecraft's noise names;
  * xoroshiro128++ with Minecraft's seeding, and improved Perlin noise:
    against independent Python restatements of the published algorithms;
  * the climate composition, its constants and the terrain spline: against
    the reference's noise sources read as text (test_biome_reference_text.py);
  * the parameter trees of every version: the generated tables equal the
    reference's table text; the version selection; the pruned nearest-leaf
    search of both node layouts equals an exhaustive search.
Biome ids against the reference's CUDA build or cubiomes: parity unpinned.
"""
import ctypes as C
import hashlib
import os
import random
import re
import struct

import pytest

from biome_restatement import ImprovedNoise, Xoroshiro128PlusPlus
from conftest import ROOT

INC = os.path.join(ROOT, "slime_amd", "csrc", "biome_data.inc")
REF_TABLES = "/root/reference/gpu_noise/tables"
TREE_IDS = list(range(7))
M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def bl():
    from oracle import pyoracle
    return pyoracle.biome_lib()


def inc_text():
    return open(INC).read()


def md5_halves(name):
    d = hashlib.md5(name.encode()).digest()
    return int.from_bytes(d[:8], "big"), int.from_bytes(d[8:], "big")


def test_noise_seeds_are_md5_of_noise_names():
    text = inc_text()
    rows = re.findall(r"\{0x([0-9a-f]{16})ull, 0x([0-9a-f]{16})ull, 0x([0-9a-f]{16})ull, "
                      r"0x([0-9a-f]{16})ull, (-?\d+), (-?\d+), (\d+),", text)
    names = [("minecraft:offset", None), ("minecraft:temperature", "minecraft:temperature_large"),
             ("minecraft:vegetation", "minecraft:vegetation_large"),
             ("minecraft:continentalness", "minecraft:continentalness_large"),
             ("minecraft:erosion", "minecraft:erosion_large"), ("minecraft:ridge", None)]
    assert len(rows) == 6
    for (lo, hi, llo, lhi, *_), (n, large) in zip(rows, names):
        assert (int(lo, 16), int(hi, 16)) == md5_halves(n)
        assert (int(llo, 16), int(lhi, 16)) == md5_halves(large or n)
    octs = re.findall(r"\{0x([0-9a-f]{16})ull, 0x([0-9a-f]{16})ull\},\s*// octave_(-?\d+)", text)
    assert len(octs) == 17
    for lo, hi, i in octs:
        assert (int(lo, 16), int(hi, 16)) == md5_halves(f"octave_{i}")


def test_xoroshiro_matches_published_algorithm(bl):
    for seed in (0, 1, 12345, M64, 0x5EED5EED5EED):
        ref = Xoroshiro128PlusPlus(seed)
        want = [ref.next_long() for _ in range(64)]
        got = (C.c_uint64 * 64)()
        bl.ob_xoro_stream(seed, 64, got)
        assert list(got) == want, seed


def test_perlin_matches_published_algorithm(bl):
    rng = random.Random(8)
    for seed in (0, 7, 0x123456789ABCDEF, M64 - 3):
        noise = ImprovedNoise(Xoroshiro128PlusPlus(seed))
        for _ in range(200):
            x, y, z = (rng.uniform(-3e6, 3e6), rng.choice([0.0, rng.uniform(-512, 512)]),
                       rng.uniform(-3e6, 3e6))
            assert bl.ob_perlin_sample_seeded(seed, x, y, z) == noise.sample(x, y, z)


def f32(v):
    return struct.unpack("<f", struct.pack("<f", v))[0]


def inc_tree(suffix):
    """(nodes, intervals) of one tree as generated into biome_data.inc."""
    text = inc_text()
    nodes = [int(h, 16) for h in re.findall(
        r"0x([0-9a-f]{16})ull", text.split(f"kTreeNodes{suffix}[", 1)[1].split("};", 1)[0])]
    param = [(int(a), int(b)) for a, b in re.findall(
        r"\{(-?\d+), (-?\d+)\}", text.split(f"kTreeParams{suffix}[", 1)[1].split("};", 1)[0])]
    return nodes, param


@pytest.mark.parametrize("tree", TREE_IDS)
def test_tree_tables_equal_reference_text(bl, tree):
    from oracle.pyoracle import BIOME_TREES
    suffix, _ = BIOME_TREES[tree]
    path = os.path.join(REF_TABLES, f"btree{suffix}.h")
    if not os.path.exists(path):
        pytest.skip("reference tree tables not present")
    ref = open(path).read()
    ref_nodes = [int(h, 16) for h in re.findall(r"\b0x([0-9A-Fa-f]{16})\b", ref.split("_nodes[]", 1)[1])]
    ref_param = [(int(a), int(b)) for a, b in re.findall(
        r"\{\s*(-?\d+)\s*,\s*(-?\d+)\s*\}", re.sub(r"//[^\n]*", "", ref.split("_param[][2]", 1)[1].split("};", 1)[0]))]
    nodes, param = inc_tree(suffix)
    assert nodes == ref_nodes and len(nodes) > 8000
    assert param == ref_param
    # the oracle reads the same tree under this index
    assert bl.ob_tree_len(tree) == len(nodes)
    assert [bl.ob_tree_node(tree, i) for i in (0, 1, len(nodes) // 2, len(nodes) - 1)] == \
        [nodes[0], nodes[1], nodes[len(nodes) // 2], nodes[-1]]
    # the reference's steps table: family 1 is 1555..1, family 0 1111..1
    steps = [int(x) for x in re.search(r"_steps\[\]\s*=\s*\{([^}]*)\}", ref).group(1).split(",") if x.strip()]
    assert steps == ([1555, 259, 43, 7, 1, 0] if bl.ob_tree_family(tree) else [1111, 111, 11, 1, 0])


def test_version_selection(bl):
    """MinecraftGPUNoise.cu:1785-1820: the first tree whose version code is
    reached; nothing below 1.18."""
    want = {1179: -1, 1180: 6, 1191: 6, 1192: 5, 1193: 5, 1194: 4, 1205: 4, 1206: 3, 1213: 3,
            1214: 2, 1215: 1, 2601: 1, 2602: 0, 9999: 0}
    for v, t in want.items():
        assert bl.ob_tree_for_version(v) == t, v


@pytest.mark.parametrize("tree", TREE_IDS)
def test_pruned_tree_search_is_nearest(bl, tree):
    rng = random.Random(11)
    ties = 0
    for i in range(3000):
        np6 = [rng.randint(-12000, 12000) for _ in range(6)]
        if i % 3 == 0:
            np6[4] = rng.randint(-2000, 12000)  # depth-like values
        a = (C.c_int32 * 6)(*np6)
        leaf = bl.ob_tree_leaf(a, tree)
        best = bl.ob_tree_leaf_exhaustive(a, tree)
        d_leaf, d_best = bl.ob_node_distance(a, leaf, tree), bl.ob_node_distance(a, best, tree)
        assert d_leaf == d_best, np6
        if leaf != best:
            ties += 1
    # equal-distance leaves exist, but the minimum is unique for most points
    assert ties < 300


def test_biome_ids_and_filter_on_a_small_world(bl):
    """Every quart of a small area gets a leaf's biome id; the filter of a
    candidate equals the union over its disk's slime chunks' quarts."""
    nodes, _ = inc_tree("262")
    leaf_ids = {(n >> 48) & 0xFF for n in nodes if n >> 56 == 0xFF}
    c = bl.ob_climate_new(1234, 0)
    try:
        seen = set()
        rng = random.Random(21)
        for _ in range(400):
            x, z = rng.randint(-60000, 60000), rng.randint(-60000, 60000)
            for y in (-16, -8, 0, 16):
                b = bl.ob_biome_at(c, x, y, z, 0)
                assert b in leaf_ids
                seen.add(b)
        assert len(seen) >= 10
        m = bl.ob_filter(c, 1234, 3, -7, 3, 0)
        assert 0 <= m <= 3
    finally:
        bl.ob_climate_free(c)


@pytest.mark.parametrize("tree", TREE_IDS)
def test_search_from_any_leaf_equals_search_from_root(bl, tree):
    """Starting the search from a known leaf (bound: its distance + 1) returns
    the root search's leaf, for random hints and for the previous layer's leaf
    of real columns."""
    from oracle.pyoracle import BIOME_TREES
    rng = random.Random(13)
    nodes, _ = inc_tree(BIOME_TREES[tree][0])
    leaves = [i for i, n in enumerate(nodes) if n >> 56 == 0xFF]
    for _ in range(2000):
        np6 = [rng.randint(-12000, 12000) for _ in range(6)]
        a = (C.c_int32 * 6)(*np6)
        cold = bl.ob_tree_leaf(a, tree)
        assert bl.ob_tree_leaf_from(a, rng.choice(leaves), tree) == cold
        assert bl.ob_tree_leaf_from(a, cold, tree) == cold
    c = bl.ob_climate_new(77, 0)
    col = (C.c_int32 * 6)()
    try:
        for _ in range(300):
            bl.ob_column(c, rng.randint(-10 ** 6, 10 ** 6), rng.randint(-10 ** 6, 10 ** 6), col)
            off = struct.unpack("<f", struct.pack("<i", col[5]))[0]
            hint = -1
            for y in [16] + list(range(0, -17, -1)):
                d = f32(1.0 - (y * 4) / 128.0 - 83.0 / 160.0 + off)
                np6 = (C.c_int32 * 6)(col[0], col[1], col[2], col[3], int(f32(10000.0 * d)), col[4])
                cold = bl.ob_tree_leaf(np6, tree)
                if hint >= 0:
                    assert bl.ob_tree_leaf_from(np6, hint, tree) == cold
                hint = cold
    finally:
        bl.ob_climate_free(c)


@pytest.mark.parametrize("tree", TREE_IDS)
def test_bounded_search_on_distance_ties(bl, tree):
    """Points with several leaves at the least distance: the search bounded by
    any of them (the kernels' start from the last leaf) returns the root
    search's leaf, the first of them in visiting order."""
    rng = random.Random(77 + tree)
    n = bl.ob_tree_len(tree)
    leaves = [i for i in range(n) if (bl.ob_tree_node(tree, i) >> 56) == 0xFF]
    ties = 0
    for it in range(400):
        np6 = [rng.randint(-12000, 12000) for _ in range(6)]
        if it % 2:
            np6[4] = rng.randint(-2000, 12000)
        a = (C.c_int32 * 6)(*np6)
        cold = bl.ob_tree_leaf(a, tree)
        d = bl.ob_node_distance(a, cold, tree)
        mins = [l for l in leaves if bl.ob_node_distance(a, l, tree) == d]
        if len(mins) > 1:
            ties += 1
        for h in mins:
            assert bl.ob_tree_leaf_from(a, h, tree) == cold, (np6, h)
    assert ties >= 1


@pytest.mark.parametrize("tree", TREE_IDS)
def test_first_descent_reaches_a_leaf(bl, tree):
    """The reference's search takes the first child at every level before it
    has found a leaf; that path ends on a leaf in every tree (so starting with
    no leaf instead of node 0 changes nothing)."""
    leaf = bl.ob_tree_first_descent(tree)
    assert leaf >= 0 and (bl.ob_tree_node(tree, leaf) >> 56) == 0xFF


def disk_mask_brute(bl, c, seed, cx, cz, ys, all_chunks, tree=0):
    """The level-range filter's mask by point queries: every sampled chunk of
    the disk, its 16 quart columns, quart y = floor(y / 4) per height."""
    from oracle.pyoracle import oracle_lib
    ol = oracle_lib()
    mask = 0
    for dx in range(-8, 9):
        for dz in range(-8, 9):
            if dx * dx + dz * dz > 68:
                continue
            x, z = cx + dx, cz + dz
            if not all_chunks and not ol.oracle_is_slime(seed, x, z):
                continue
            for q in range(16):
                qx, qz = 4 * x + q // 4, 4 * z + q % 4
                for y in ys:
                    b = bl.ob_biome_at(c, qx, y >> 2, qz, tree)
                    mask |= (1 if b == 183 else 0) | (2 if b == 14 else 0)
    return mask


def deep_dark_site(bl, c, rng, yq=-10, tree=0):
    """A candidate whose disk holds a Deep Dark quart at quart y yq (the chunk
    3 east and 2 north of the center)."""
    for _ in range(20000):
        qx, qz = rng.randint(-2_000_000, 2_000_000), rng.randint(-2_000_000, 2_000_000)
        if bl.ob_biome_at(c, qx, yq, qz, tree) == 183:
            return (qx >> 2) - 3, (qz >> 2) + 2
    return None


def test_level_filter_reduces_to_the_reference_set(bl):
    """Heights -64..0 (any step up to 4, any phase within the quart) with Deep
    Dark alone, and 64 alone with Mushroom Fields, over slime chunks, give the
    reference filter's masks."""
    seed = 0x0DDBA11CAFE5EED5
    rng = random.Random(71)
    c = bl.ob_climate_new(seed, 0)
    try:
        cands = [(rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000))
                 for _ in range(6)]
        site = deep_dark_site(bl, c, rng)
        assert site is not None
        cands.append(site)
        found = 0
        for x, z in cands:
            m1 = bl.ob_filter(c, seed, x, z, 1, 0)
            found |= m1
            for lo, hi, step in ((-64, 0, 4), (-64, 3, 1), (-61, 3, 4), (-63, 2, 2)):
                assert bl.ob_filter_levels(c, seed, x, z, 1, lo, hi, step, 0) == m1, (lo, hi, step)
            m2 = bl.ob_filter(c, seed, x, z, 2, 0)
            for lo, hi, step in ((64, 64, 1), (64, 67, 1), (67, 70, 8)):
                assert bl.ob_filter_levels(c, seed, x, z, 2, lo, hi, step, 0) == m2, (lo, hi, step)
        # the planted chunk need not be a slime chunk; with every chunk it counts
        x, z = site
        assert bl.ob_filter_levels(c, seed, x, z, 1 | 4, -64, 0, 4, 0) == 1
    finally:
        bl.ob_climate_free(c)


@pytest.mark.parametrize("all_chunks", [False, True])
def test_level_filter_equals_point_queries(bl, all_chunks):
    """The level-range mask equals the union over point queries of the disk's
    sampled chunks, at heights that fall in the middle of quarts and below 0."""
    seed = -3141592653589793238
    rng = random.Random(72 + all_chunks)
    c = bl.ob_climate_new(seed & ((1 << 64) - 1), 0)
    try:
        site = deep_dark_site(bl, c, rng, yq=-10)
        cands = [site, (rng.randint(-100_000, 100_000), rng.randint(-100_000, 100_000))]
        ys = list(range(-38, 40, 19))  # -38, -19, 0, 19, 38: quarts -10, -5, 0, 4, 9
        flags = 3 | (4 if all_chunks else 0)
        masks = []
        for x, z in cands:
            got = bl.ob_filter_levels(c, seed, x, z, flags, ys[0], ys[-1], 19, 0)
            assert got == disk_mask_brute(bl, c, seed, x, z, ys, all_chunks)
            masks.append(got)
        if all_chunks:
            assert masks[0] & 1  # the planted quart lies at quart y -10 = floor(-38 / 4)
    finally:
        bl.ob_climate_free(c)
