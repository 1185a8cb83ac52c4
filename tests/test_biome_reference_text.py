"""The biome checker (oracle/biome_oracle.c) and the data it shares with the
kernels (slime_amd/csrc/biome_data.inc) against the reference's noise sources
read as text (gpu_noise/MinecraftGPUNoise.cu, gpu_noise/spline_generated.cuh):

  * every constant of the climate composition: per-noise seed halves (normal
    and large biomes), first octave, octave count and amplitudes; the octave
    seeds; the lacunarity, persistence and double-Perlin amplitude tables;
  * the composition itself: climate doubles, quantized parameters, depth and
    spline offset of quart columns equal an independent restatement
    (tests/biome_restatement.py) built from those constants and the
    published Perlin/xoroshiro algorithms, bit for bit, for normal and large
    biomes;
  * the terrain spline: equal, bit for bit, to the reference's generated
    spline source transliterated statement by statement into numpy
    float32/float64 (20,000 vectors, every knot and its neighbours);
  * planted mistakes in the data (a child index, a derivative, a knot, an
    amplitude, a seed half, an octave count) make these checks fail.

Skipped where the reference tree is absent (these are CPU tests).
"""
import ctypes as C
import os
import random
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT
import biome_restatement as R

pytestmark = pytest.mark.skipif(not R.ref_available(), reason="reference noise sources absent")

INC = os.path.join(ROOT, "slime_amd", "csrc", "biome_data.inc")
SLOTS = ["SHIFT", "TEMPERATURE", "HUMIDITY", "CONTINENTALNESS", "EROSION", "WEIRDNESS"]


@pytest.fixture(scope="module")
def consts():
    return R.NoiseConstants()


@pytest.fixture(scope="module")
def spline():
    return R.spline_functions()[1]


@pytest.fixture(scope="module")
def bl():
    from oracle import pyoracle
    return pyoracle.biome_lib()


def inc_noise_rows(text):
    """kBiomeNoises of biome_data.inc: per noise (lo, hi, large lo, large hi,
    first octave, large first octave, length, amplitudes)."""
    body = text.split("kBiomeNoises[6] = {", 1)[1].split("};", 1)[0]
    rows = []
    for m in re.finditer(r"\{0x([0-9a-f]{16})ull, 0x([0-9a-f]{16})ull, 0x([0-9a-f]{16})ull, "
                         r"0x([0-9a-f]{16})ull, (-?\d+), (-?\d+), (\d+), \{([^}]*)\}\}", body):
        g = m.groups()
        rows.append((int(g[0], 16), int(g[1], 16), int(g[2], 16), int(g[3], 16), int(g[4]),
                     int(g[5]), int(g[6]), [float(a) for a in g[7].split(",")]))
    assert len(rows) == 6
    return rows


def noise_table_mismatches(rows, consts):
    bad = []
    for slot, (lo, hi, llo, lhi, omin, lomin, length, amps) in zip(SLOTS, rows):
        d = consts.noises[slot]
        want_amps = d["amps"] + [0.0] * (9 - len(d["amps"]))
        for name, got, want in (("lo", lo, d["lo"]), ("hi", hi, d["hi"]),
                                ("large_lo", llo, d["large_lo"]), ("large_hi", lhi, d["large_hi"]),
                                ("omin", omin, d["omin"]), ("large_omin", lomin, d["large_omin"]),
                                ("len", length, d["len"]), ("amps", amps, want_amps)):
            if got != want:
                bad.append((slot, name, got, want))
    return bad


def test_noise_table_equals_reference_text(consts):
    rows = inc_noise_rows(open(INC).read())
    assert noise_table_mismatches(rows, consts) == []
    octs = re.findall(r"\{0x([0-9a-f]{16})ull, 0x([0-9a-f]{16})ull\},\s*// octave_(-?\d+)",
                      open(INC).read())
    assert [(int(a, 16), int(b, 16)) for a, b, _ in octs] == consts.octave_seeds


@pytest.mark.parametrize("slot", range(6))
def test_planted_noise_table_mistakes_are_caught(consts, slot):
    rows = inc_noise_rows(open(INC).read())
    lo, hi, llo, lhi, omin, lomin, length, amps = rows[slot]
    plants = [(lo ^ 1, hi, llo, lhi, omin, lomin, length, amps),
              (lo, hi, llo, lhi ^ (1 << 63), omin, lomin, length, amps),
              (lo, hi, llo, lhi, omin - 1, lomin, length, amps),
              (lo, hi, llo, lhi, omin, lomin + 1, length, amps),
              (lo, hi, llo, lhi, omin, lomin, length - 1, amps)]
    for i in range(length):
        plants.append((lo, hi, llo, lhi, omin, lomin, length,
                       amps[:i] + [amps[i] + 0.5] + amps[i + 1:]))
    for p in plants:
        bad = list(rows)
        bad[slot] = p
        assert noise_table_mismatches(bad, consts), p


def test_octave_tables_are_the_oracles_closed_forms(consts):
    """The oracle computes lacunarity as 2^omin, persistence as
    2^(n-1)/(2^n-1) and the double-Perlin amplitude as 5n/(3(n+1)); the
    reference tabulates them (MinecraftGPUNoise.cu:227-253)."""
    for i in range(17):
        assert consts.lacuna[i] == 2.0 ** -i
    for n in range(1, 17):
        assert consts.persist[n] == float(1 << (n - 1)) / float((1 << n) - 1)
        assert consts.double_amp[n] == float(5 * n) / float(3 * (n + 1))
    assert consts.persist[0] == 0.0 and consts.double_amp[0] == 0.0


def column_points(rng, n):
    pts = [(0, 0), (-1, 1), (7_499_999, -7_500_000), (-7_500_000, 7_499_999)]
    while len(pts) < n:
        pts.append((rng.randint(-7_500_000, 7_500_000), rng.randint(-7_500_000, 7_500_000)))
    return pts


@pytest.mark.parametrize("seed,large", [(0, 0), (1234, 0), (-7995527694508729151, 0),
                                        (0x5EED000000000003, 1), (-42, 1)])
def test_climate_composition_equals_reference(bl, consts, spline, seed, large):
    rng = random.Random(seed ^ large)
    ref = R.RefClimate(consts, seed, bool(large))
    c = bl.ob_climate_new(seed & R.M64, large)
    try:
        out5 = (C.c_double * 5)()
        col = (C.c_int32 * 6)()
        out6 = (C.c_int32 * 6)()
        for x, z in column_points(rng, 12):
            bl.ob_climate5(c, float(x), float(z), out5)
            assert list(out5) == ref.climate5(float(x), float(z)), (x, z)
            bl.ob_column(c, x, z, col)
            np6, off = ref.climate6(spline, x, 0, z)
            assert struct.pack("<i", col[5]) == struct.pack("<f", off), (x, z)
            assert list(col[:5]) == [np6[0], np6[1], np6[2], np6[3], np6[5]], (x, z)
            for y in (-16, -3, 0, 16):
                bl.ob_climate6(c, x, y, z, out6)
                assert list(out6) == ref.climate6(spline, x, y, z)[0], (x, y, z)
        # fractional points of the double-valued query
        for _ in range(4):
            x, z = rng.uniform(-3e7, 3e7), rng.uniform(-3e7, 3e7)
            bl.ob_climate5(c, x, z, out5)
            assert list(out5) == ref.climate5(x, z)
    finally:
        bl.ob_climate_free(c)


def reference_knots():
    text = open(os.path.join(R.REF_NOISE, "spline_generated.cuh")).read()
    return sorted({float(np.float32(k)) for k in re.findall(r"if \(f <= (-?\d+\.\d*(?:e[+-]?\d+)?)f\)", text)})


def spline_vectors(n, seed=3):
    rng = random.Random(seed)
    knots = reference_knots()
    out = []
    for i in range(n):
        v = [rng.uniform(-1.3, 1.3) for _ in range(4)]
        if i % 4 == 0:  # on a knot, one float below, one float above
            k = np.float32(rng.choice(knots))
            step = rng.choice([0, -1, 1])
            v[rng.randrange(4)] = float(np.nextafter(k, np.float32(step * 10), dtype=np.float32)
                                        if step else k)
        if i % 2:  # peaks and valleys as the climate derives it from weirdness
            w = np.float32(v[3])
            v[2] = float(-(abs(abs(w) - np.float32(0.6666667)) - np.float32(0.33333334))
                         * np.float32(3.0))
        out.append(np.array(v, dtype=np.float32))
    return out


def spline_mismatches(lib, spline, vectors):
    bad = 0
    for v in vectors:
        a = (C.c_float * 4)(*[float(x) for x in v])
        if struct.pack("<f", lib.ob_spline(a)) != struct.pack("<f", spline(v)):
            bad += 1
    return bad


def test_spline_equals_reference_text(bl, spline):
    vectors = spline_vectors(20000)
    assert spline_mismatches(bl, spline, vectors) == 0
    # every knot of every spline, exactly and one float either side, on each coordinate
    vs = []
    for k in reference_knots():
        for kk in (np.float32(k), np.nextafter(np.float32(k), np.float32(9), dtype=np.float32),
                   np.nextafter(np.float32(k), np.float32(-9), dtype=np.float32)):
            for i in range(4):
                v = np.array([0.1, -0.2, 0.3, 0.45], dtype=np.float32)
                v[i] = kk
                vs.append(v)
    assert spline_mismatches(bl, spline, vs) == 0


# -- planted mistakes in the shared data: the checks above must fail ---------------

def _spline_rows(text):
    return [m for m in re.finditer(r"\{0x([0-9a-f]{8})u, 0x([0-9a-f]{8})u, 0x([0-9a-f]{8})u, (-?\d+)\}",
                                   text)]


def plant(text, kind):
    rows = _spline_rows(text)
    if kind == "child":          # the first point with a child spline names the next spline
        m = next(r for r in rows if int(r.group(4)) >= 0)
        new = f"{{0x{m.group(1)}u, 0x{m.group(2)}u, 0x{m.group(3)}u, {int(m.group(4)) + 1}}}"
    elif kind == "derivative":   # a derivative and a value swapped
        m = next(r for r in rows if int(r.group(4)) < 0 and r.group(2) != r.group(3)
                 and r.group(2) != "00000000")
        new = f"{{0x{m.group(1)}u, 0x{m.group(3)}u, 0x{m.group(2)}u, {m.group(4)}}}"
    elif kind == "knot":         # a knot one float off
        m = rows[len(rows) // 2]
        new = f"{{0x{int(m.group(1), 16) + 1:08x}u, 0x{m.group(2)}u, 0x{m.group(3)}u, {m.group(4)}}}"
    elif kind == "amplitude":    # continentalness octave 3 amplitude 2.0 -> 1.0
        old = "{1.0, 1.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0, 1.0}"
        assert old in text
        return text.replace(old, "{1.0, 1.0, 2.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0}", 1)
    else:
        raise ValueError(kind)
    return text[:m.start()] + new + text[m.end():]


@pytest.fixture(scope="module")
def planted_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("planted")


def build_planted(tmp, kind):
    """The oracle compiled against biome_data.inc with one planted mistake."""
    d = tmp / kind
    (d / "oracle").mkdir(parents=True)
    (d / "slime_amd" / "csrc").mkdir(parents=True)
    shutil.copy(os.path.join(ROOT, "oracle", "biome_oracle.c"), d / "oracle")
    text = open(INC).read()
    planted = plant(text, kind)
    assert planted != text
    (d / "slime_amd" / "csrc" / "biome_data.inc").write_text(planted)
    so = d / "oracle" / "libbiome_planted.so"
    subprocess.run(["gcc", "-std=c11", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-o", str(so),
                    str(d / "oracle" / "biome_oracle.c"), "-lm"], check=True)
    from oracle import pyoracle
    return pyoracle.load_biome_lib(str(so))


@pytest.mark.parametrize("kind", ["child", "derivative", "knot"])
def test_planted_spline_mistakes_are_caught(planted_dir, spline, kind):
    lib = build_planted(planted_dir, kind)
    vectors = spline_vectors(4000, seed=5)
    vs = []
    for k in reference_knots():
        for i in range(4):
            v = np.array([0.1, -0.2, 0.3, 0.45], dtype=np.float32)
            v[i] = np.float32(k)
            vs.append(v)
    assert spline_mismatches(lib, spline, vectors + vs) > 0


def test_planted_amplitude_mistake_is_caught(planted_dir, consts):
    lib = build_planted(planted_dir, "amplitude")
    ref = R.RefClimate(consts, 1234, False)
    c = lib.ob_climate_new(1234, 0)
    try:
        out5 = (C.c_double * 5)()
        lib.ob_climate5(c, 1000.0, -2000.0, out5)
        assert list(out5) != ref.climate5(1000.0, -2000.0)
    finally:
        lib.ob_climate_free(c)


def test_filter_equals_reference_sampling(bl, consts, spline):
    """The oracle's candidate filter (ob_filter) equals the reference kernel's
    sampling set read from its text (the disk's slime chunks, 16 quart columns
    each, Deep Dark at quart y -16..0, Mushroom Fields at 16), with every
    biome from the reference-text climate above and the exhaustively pinned
    tree search; on candidates with and without a Deep Dark quart."""
    from oracle import pyoracle
    ol = pyoracle.oracle_lib()
    samp = R.FilterSampling(consts.text)
    assert (samp.side, samp.half, samp.r2, samp.quarts, samp.layers) == (17, 8, 68, 16, 17)
    seed = -7995527694508729151
    ref = R.RefClimate(consts, seed, False)
    c = bl.ob_climate_new(seed & R.M64, 0)
    rng = random.Random(91)
    found = {0: None, 1: None}
    try:
        for _ in range(200):
            x, z = rng.randint(-1_800_000, 1_800_000), rng.randint(-1_800_000, 1_800_000)
            m = bl.ob_filter(c, seed, x, z, 3, 0)
            if found[m & 1] is None:
                found[m & 1] = (x, z, m)
            if None not in found.values():
                break
        assert None not in found.values()
        columns = {}

        def biome_of(qx, qy, qz):
            if (qx, qz) not in columns:
                columns[(qx, qz)] = ref.column(spline, qx, qz)
            np6 = (C.c_int32 * 6)(*ref.params(columns[(qx, qz)], qy))
            return bl.ob_leaf_biome(bl.ob_tree_leaf(np6, 0), 0)

        for x, z, m in found.values():
            for flags in (1, 2, 3):
                want = samp.mask(lambda s, a, b: bool(ol.oracle_is_slime(s, a, b)), biome_of,
                                 seed, x, z, flags)
                assert bl.ob_filter(c, seed, x, z, flags, 0) == want, (x, z, flags)
            assert want == m
    finally:
        bl.ob_climate_free(c)
