"""GPU parity of the biome filter (mcgpu_modern_ctx_*) against the CPU
restatement in oracle/biome_oracle.c: the per-column climate, biome ids at
quart points, and filter masks of slime-cluster candidates, bit-exact.
(Biome ids against the reference's CUDA build: parity unpinned; see
tests/test_biome_oracle.py for what pins the restatement.)"""
import ctypes as C
import random

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bl():
    from oracle import pyoracle
    return pyoracle.biome_lib()


def oracle_columns(bl, clim, xs, zs):
    out = []
    buf = (C.c_int32 * 6)()
    for x, z in zip(xs, zs):
        bl.ob_column(clim, x, z, buf)
        out.append(tuple(buf))
    return out


@pytest.mark.parametrize("large", [False, True])
def test_columns_and_queries_match_oracle(gpu, bl, large):
    import slime_amd
    seed = 0x6D696E6563726166 if large else -4172144997902289642
    rng = random.Random(5 + large)
    n = 3000
    xs = [rng.randint(-7_500_000, 7_500_000) for _ in range(n)]
    zs = [rng.randint(-7_500_000, 7_500_000) for _ in range(n)]
    ys = [rng.randint(-20, 20) for _ in range(n)]
    clim = bl.ob_climate_new(seed & ((1 << 64) - 1), 1 if large else 0)
    try:
        with slime_amd.BiomeContext(seed, capacity=n, large_biomes=large) as ctx:
            got_cols = ctx.columns(xs, zs)
            want_cols = oracle_columns(bl, clim, xs, zs)
            assert got_cols == want_cols
            got = ctx.query(xs, ys, zs)
            want = [bl.ob_biome_at(clim, x, y, z) for x, y, z in zip(xs, ys, zs)]
            assert got == want
            assert len(set(got)) >= 10
    finally:
        bl.ob_climate_free(clim)


def find_biome_chunk(bl, clim, seed, biome, ys, rng, tries=20000):
    """A slime chunk holding a quart of the biome at one of the layers ys."""
    from oracle import pyoracle
    ol = pyoracle.oracle_lib()
    for _ in range(tries):
        qx, qz = rng.randint(-4_000_000, 4_000_000), rng.randint(-4_000_000, 4_000_000)
        if any(bl.ob_biome_at(clim, qx, y, qz) == biome for y in ys):
            cx, cz = qx >> 2, qz >> 2
            if ol.oracle_is_slime(seed, cx, cz):
                return cx, cz
    return None


def test_filter_matches_oracle(gpu, bl):
    import slime_amd
    seed = 0x1F2E3D4C5B6A7988
    rng = random.Random(9)
    clim = bl.ob_climate_new(seed, 0)
    try:
        cands = [(rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000))
                 for _ in range(20)]
        planted = []
        dd = find_biome_chunk(bl, clim, seed, 183, range(-16, 1, 4), rng)
        mf = find_biome_chunk(bl, clim, seed, 14, [16], rng, tries=60000)
        for hit in (dd, mf):
            if hit:
                planted.append((hit[0] + 3, hit[1] - 2))  # the chunk lies in this disk
        cands += planted
        print("planted", dd, mf, flush=True)
        cx, cz = [c[0] for c in cands], [c[1] for c in cands]
        with slime_amd.BiomeContext(seed, capacity=len(cands)) as ctx:
            for flags in (3, 1, 2):
                got = ctx.filter(cx, cz, flags)
                want = [bl.ob_filter(clim, seed, x, z, flags) for x, z in cands]
                print("flags", flags, got, flush=True)
                assert got == want
            if dd:
                assert ctx.filter([cx[20]], [cz[20]], 1)[0] & 1
    finally:
        bl.ob_climate_free(clim)


def test_filter_rejects_bad_arguments(gpu):
    import slime_amd
    L = slime_amd.lib()
    assert L.mcgpu_modern_ctx_create(2601, 1, 16) is None  # older trees are not carried
    assert L.mcgpu_modern_ctx_create(2602, 1, 0) is None
    with slime_amd.BiomeContext(1, capacity=4) as ctx:
        a = (C.c_int32 * 8)()
        assert L.mcgpu_modern_ctx_filter_candidates(ctx._h, a, a, 5, 3, a) == -1  # > capacity
        assert L.mcgpu_modern_ctx_filter_candidates(ctx._h, a, a, 2, 4, a) == -1  # bad flags
        assert L.mcgpu_modern_ctx_filter_candidates(ctx._h, a, a, 2, 0, a) == -1


def test_filter_multi_equals_per_world_contexts(gpu, bl):
    """Candidates of several worlds in one launch: the same masks as each
    world's own context, and as the oracle."""
    import slime_amd
    rng = random.Random(31)
    seeds = [rng.getrandbits(64) for _ in range(5)]
    cands = [[(rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000))
              for _ in range(rng.randint(0, 9))] for _ in seeds]
    masks, ms = slime_amd.filter_multi(seeds, cands, 3)
    assert ms > 0
    for seed, cs, got in zip(seeds, cands, masks):
        if not cs:
            assert got == []
            continue
        with slime_amd.BiomeContext(seed, capacity=len(cs)) as ctx:
            assert got == ctx.filter([c[0] for c in cs], [c[1] for c in cs], 3)
        clim = bl.ob_climate_new(seed, 0)
        try:
            assert got == [bl.ob_filter(clim, seed - (1 << 64) if seed >> 63 else seed, x, z, 3)
                           for x, z in cs]
        finally:
            bl.ob_climate_free(clim)
