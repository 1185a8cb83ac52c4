"""GPU parity of the biome filter and point queries (mcgpu_*) against the CPU
restatement in oracle/biome_oracle.c: the per-column climate, biome ids at
quart points for every version's tree, filter masks of slime-cluster
candidates, and the noise / climate primitives, bit-exact.
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
            want = [bl.ob_biome_at(clim, x, y, z, 0) for x, y, z in zip(xs, ys, zs)]
            assert got == want
            assert len(set(got)) >= 10
    finally:
        bl.ob_climate_free(clim)


def find_biome_chunk(bl, clim, seed, biome, ys, rng, tries=20000, tree=0):
    """A slime chunk holding a quart of the biome at one of the layers ys."""
    from oracle import pyoracle
    ol = pyoracle.oracle_lib()
    for _ in range(tries):
        qx, qz = rng.randint(-4_000_000, 4_000_000), rng.randint(-4_000_000, 4_000_000)
        if any(bl.ob_biome_at(clim, qx, y, qz, tree) == biome for y in ys):
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
                want = [bl.ob_filter(clim, seed, x, z, flags, 0) for x, z in cands]
                print("flags", flags, got, flush=True)
                assert got == want
            if dd:
                assert ctx.filter([cx[20]], [cz[20]], 1)[0] & 1
    finally:
        bl.ob_climate_free(clim)


def test_filter_rejects_bad_arguments(gpu):
    import slime_amd
    L = slime_amd.lib()
    assert L.mcgpu_modern_ctx_create(1179, 1, 16) is None  # before 1.18
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
            assert got == [bl.ob_filter(clim, seed - (1 << 64) if seed >> 63 else seed, x, z, 3, 0)
                           for x, z in cs]
        finally:
            bl.ob_climate_free(clim)


VERSIONS = [(2602, 0), (1215, 1), (1214, 2), (1206, 3), (1194, 4), (1192, 5), (1180, 6)]


@pytest.mark.parametrize("version,tree", VERSIONS)
def test_every_version_tree(gpu, bl, version, tree):
    """Queries and filter masks with each version's tree (both node layouts)
    against the oracle; 1.18 has no Deep Dark leaf."""
    import slime_amd
    seed = 0x0DDBA11 + version
    rng = random.Random(version)
    n = 1500
    xs = [rng.randint(-3_000_000, 3_000_000) for _ in range(n)]
    zs = [rng.randint(-3_000_000, 3_000_000) for _ in range(n)]
    ys = [rng.randint(-20, 20) for _ in range(n)]
    clim = bl.ob_climate_new(seed, 0)
    try:
        with slime_amd.BiomeContext(seed, capacity=n, version=version) as ctx:
            got = ctx.query(xs, ys, zs)
            assert got == [bl.ob_biome_at(clim, x, y, z, tree) for x, y, z in zip(xs, ys, zs)]
            cands = [(rng.randint(-500_000, 500_000), rng.randint(-500_000, 500_000))
                     for _ in range(12)]
            mf = find_biome_chunk(bl, clim, seed, 14, [16], rng, tries=40000, tree=tree)
            if mf:
                cands.append((mf[0] - 1, mf[1] + 4))
            cx, cz = [c[0] for c in cands], [c[1] for c in cands]
            print("version", version, "planted", mf, flush=True)
            got = ctx.filter(cx, cz, 3)
            assert got == [bl.ob_filter(clim, seed, x, z, 3, tree) for x, z in cands]
            if mf:
                assert got[-1] & 2
            if version < 1190:
                assert all(m & 1 == 0 for m in got)
        masks, _ = slime_amd.filter_multi([seed], [cands], 3, version=version)
        assert masks == [got]
    finally:
        bl.ob_climate_free(clim)


def quant(v):
    import numpy as np
    return int(np.float32(10000.0) * np.float32(v))


def test_noise_primitives(gpu, bl):
    """xoroshiro128++ streams and Perlin noise (per-query seeds and one fixed
    seed) against the restatement."""
    import slime_amd
    rng = random.Random(44)
    seeds = [rng.getrandbits(64) for _ in range(300)] + [0, (1 << 64) - 1]
    got = slime_amd.xoroshiro_nextlong(seeds, 9)
    buf = (C.c_uint64 * 9)()
    for s_, g in zip(seeds, got):
        bl.ob_xoro_stream(s_, 9, buf)
        assert g == list(buf)
    n = 2000
    xs = [rng.uniform(-3e6, 3e6) for _ in range(n)]
    ys = [rng.choice([0.0, rng.uniform(-600, 600)]) for _ in range(n)]
    zs = [rng.uniform(-3e6, 3e6) for _ in range(n)]
    ps = [rng.getrandbits(64) for _ in range(n)]
    assert slime_amd.perlin_sample(ps, xs, ys, zs) == \
        [bl.ob_perlin_sample_seeded(p, x, y, z) for p, x, y, z in zip(ps, xs, ys, zs)]
    fixed = 0xC0FFEE123
    assert slime_amd.perlin_sample(fixed, xs, ys, zs) == \
        [bl.ob_perlin_sample_seeded(fixed, x, y, z) for x, y, z in zip(xs, ys, zs)]


@pytest.mark.parametrize("large", [False, True])
def test_climate_point_queries(gpu, bl, large):
    """climate5 (doubles, quantized, int32 points), climate6 (with depth; large
    biomes through the _ex form), the climate context, the continentalness
    noise and the 1.20 biome queries, against the restatement."""
    import slime_amd
    seed = -2 if large else 0x5EED1E55
    rng = random.Random(7 + large)
    n = 1200
    fx = [rng.uniform(-2e6, 2e6) for _ in range(n)]
    fz = [rng.uniform(-2e6, 2e6) for _ in range(n)]
    ix = [rng.randint(-2_000_000, 2_000_000) for _ in range(n)]
    iz = [rng.randint(-2_000_000, 2_000_000) for _ in range(n)]
    iy = [rng.randint(-20, 90) for _ in range(n)]
    clim = bl.ob_climate_new(seed & ((1 << 64) - 1), 1 if large else 0)
    o5 = (C.c_double * 5)()
    o6 = (C.c_int32 * 6)()
    try:
        got6 = slime_amd.climate6(seed, ix, iy, iz, large_biomes=large)
        want6 = []
        for x, y, z in zip(ix, iy, iz):
            bl.ob_climate6(clim, x, y, z, o6)
            want6.append(tuple(o6))
        assert got6 == want6
        if large:
            return
        want5 = []
        for x, z in zip(fx, fz):
            bl.ob_climate5(clim, x, z, o5)
            want5.append(tuple(o5))
        assert slime_amd.climate5(seed, fx, fz) == want5
        assert slime_amd.climate5(seed, fx, fz, quantized=True) == \
            [tuple(quant(v) for v in w) for w in want5]
        wi = []
        for x, z in zip(ix, iz):
            bl.ob_climate5(clim, float(x), float(z), o5)
            wi.append(tuple(quant(v) for v in o5))
        assert slime_amd.climate5(seed, ix, iz, integer=True) == wi
        L = slime_amd.lib()
        h = L.mcgpu_climate_ctx_create(seed, n)
        assert h
        try:
            out = (C.c_int32 * (5 * n))()
            assert L.mcgpu_climate_ctx_query_i32(h, (C.c_int32 * n)(*ix), (C.c_int32 * n)(*iz), n,
                                                 out) == 0
            assert [tuple(out[5 * i:5 * i + 5]) for i in range(n)] == wi
            assert L.mcgpu_climate_ctx_query_i32(h, (C.c_int32 * n)(*ix), (C.c_int32 * n)(*iz),
                                                 n + 1, out) == -1
        finally:
            L.mcgpu_climate_ctx_destroy(h)
        assert slime_amd.continentalness(seed, fx, fz) == \
            [bl.ob_continentalness(clim, x, z) for x, z in zip(fx, fz)]
        want_b = [bl.ob_biome_at(clim, x, y, z, 3) for x, y, z in zip(ix, iy, iz)]
        assert slime_amd.biome20(seed, ix, iy, iz) == want_b
        h = L.mcgpu_biome20_ctx_create(seed, n)
        assert h
        try:
            out = (C.c_int32 * n)()
            assert L.mcgpu_biome20_ctx_query(h, (C.c_int32 * n)(*ix), (C.c_int32 * n)(*iy),
                                             (C.c_int32 * n)(*iz), n, out) == 0
            assert list(out) == want_b
        finally:
            L.mcgpu_biome20_ctx_destroy(h)
    finally:
        bl.ob_climate_free(clim)


@pytest.mark.parametrize("version,tree", [(2602, 0), (1206, 3), (1192, 5)])
def test_filter_many_candidates(gpu, bl, version, tree):
    """A few hundred candidates per tree: masks equal the oracle's searches
    from the root (the kernel bounds each search by leaves it found before)."""
    import slime_amd
    seed = 0x5EED0000 + version
    rng = random.Random(3 * version)
    cands = [(rng.randint(-600_000, 600_000), rng.randint(-600_000, 600_000)) for _ in range(250)]
    clim = bl.ob_climate_new(seed, 0)
    try:
        with slime_amd.BiomeContext(seed, capacity=len(cands), version=version) as ctx:
            got = ctx.filter([c[0] for c in cands], [c[1] for c in cands], 3)
        want = [bl.ob_filter(clim, seed, x, z, 3, tree) for x, z in cands]
        assert got == want
        print("version", version, "deep dark", sum(m & 1 for m in got), "mushroom",
              sum(m >> 1 & 1 for m in got), flush=True)
    finally:
        bl.ob_climate_free(clim)


def deep_dark_site(bl, clim, rng, yq=-10, tree=0):
    """A candidate whose disk holds a Deep Dark quart at quart y yq (in the
    chunk 3 east and 2 north of the center, slime or not)."""
    for _ in range(20000):
        qx, qz = rng.randint(-2_000_000, 2_000_000), rng.randint(-2_000_000, 2_000_000)
        if bl.ob_biome_at(clim, qx, yq, qz, tree) == 183:
            return (qx >> 2) - 3, (qz >> 2) + 2
    return None


@pytest.mark.parametrize("version,tree", [(2602, 0), (1180, 6)])
def test_level_filter_matches_oracle(gpu, bl, version, tree):
    """The level-range filter (the wider sweep y -64..320 step 4, both biomes
    at every height) over slime chunks and over every chunk of the disk,
    against the restatement."""
    import slime_amd
    seed = 0x51DE5EED0000002A + tree
    rng = random.Random(40 + tree)
    clim = bl.ob_climate_new(seed, 0)
    try:
        cands = [(rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000))
                 for _ in range(6)]
        for yq in (-10, -2):
            site = deep_dark_site(bl, clim, rng, yq, tree)  # none before 1.19
            if site:
                cands.append(site)
        n_dd = len(cands) - 6
        mf = find_biome_chunk(bl, clim, seed, 14, [16, 40], rng, tries=60000, tree=tree)
        if mf:
            cands.append((mf[0] - 1, mf[1] + 4))
        print("candidates", cands, flush=True)
        cx, cz = [c[0] for c in cands], [c[1] for c in cands]
        with slime_amd.BiomeContext(seed, capacity=len(cands), version=version) as ctx:
            for flags, levels in ((3, (-64, 320, 4)), (3 | 4, (-64, 320, 16)), (1 | 4, (-64, 0, 4)),
                                  (2, (-70, 330, 5))):
                got = ctx.filter_levels(cx, cz, *levels, flags=flags)
                want = [bl.ob_filter_levels(clim, seed, x, z, flags, *levels, tree)
                        for x, z in cands]
                print("flags", flags, levels, got, flush=True)
                assert got == want
            got = ctx.filter_levels(cx, cz, -64, 0, 4, flags=1 | 4)
            assert all(m & 1 for m in got[6:6 + n_dd])
    finally:
        bl.ob_climate_free(clim)


def test_level_filter_reduces_to_the_reference_filter(gpu):
    """Over slime chunks, Deep Dark at heights -64..0 (steps up to 4) and
    Mushroom Fields at 64 give the reference filter's masks, candidate by
    candidate, on many candidates."""
    import slime_amd
    seed = 0x7E57AB1E
    rng = random.Random(44)
    n = 600
    cx = [rng.randint(-1_800_000, 1_800_000) for _ in range(n)]
    cz = [rng.randint(-1_800_000, 1_800_000) for _ in range(n)]
    with slime_amd.BiomeContext(seed, capacity=n) as ctx:
        m1 = ctx.filter(cx, cz, 1)
        assert sum(m1) > 0
        for levels in ((-64, 0, 4), (-64, 3, 1), (-61, 3, 4), (-63, 2, 2)):
            assert ctx.filter_levels(cx, cz, *levels, flags=1) == m1, levels
        m2 = ctx.filter(cx, cz, 2)
        for levels in ((64, 64, 1), (64, 67, 1)):
            assert ctx.filter_levels(cx, cz, *levels, flags=2) == m2, levels
        # both biomes over the range, every chunk: a superset of both
        m3 = ctx.filter_levels(cx, cz, -64, 320, 4, flags=3 | 4)
        assert all((a | b) & ~c == 0 for a, b, c in zip(m1, m2, m3))


def test_level_filter_multi_equals_per_world_contexts(gpu):
    import slime_amd
    rng = random.Random(45)
    seeds = [rng.getrandbits(64) for _ in range(4)]
    cands = [[(rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000))
              for _ in range(rng.randint(0, 40))] for _ in seeds]
    levels = (-64, 320, 4)
    masks, ms = slime_amd.filter_multi(seeds, cands, 3 | slime_amd.ALL_CHUNKS, levels=levels)
    assert ms > 0
    for seed, cs, got in zip(seeds, cands, masks):
        if not cs:
            assert got == []
            continue
        with slime_amd.BiomeContext(seed, capacity=len(cs)) as ctx:
            assert got == ctx.filter_levels([c[0] for c in cs], [c[1] for c in cs], *levels,
                                            flags=3 | slime_amd.ALL_CHUNKS)


def test_level_filter_rejects_bad_arguments(gpu):
    import slime_amd
    L = slime_amd.lib()
    with slime_amd.BiomeContext(1, capacity=4) as ctx:
        a = (C.c_int32 * 8)()
        f = L.mcgpu_modern_ctx_filter_candidates_levels
        assert f(ctx._h, a, a, 2, 3, -64, 320, 4, a) == 0
        assert f(ctx._h, a, a, 2, 3, -64, 320, 0, a) == -1  # step
        assert f(ctx._h, a, a, 2, 3, 10, 9, 1, a) == -1  # empty range
        assert f(ctx._h, a, a, 2, 3, 0, 4096, 1, a) == -1  # 4097 heights
        assert f(ctx._h, a, a, 2, 3, 0, 4095, 1, a) == 0
        assert f(ctx._h, a, a, 2, 4, -64, 0, 4, a) == -1  # no biome
        assert f(ctx._h, a, a, 2, 8 | 1, -64, 0, 4, a) == -1
        assert f(ctx._h, a, a, 5, 3, -64, 0, 4, a) == -1  # > capacity
        assert f(ctx._h, a, a, 1, 3, -(1 << 24) - 1, 0, 1 << 20, a) == -1
