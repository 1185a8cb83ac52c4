"""Independent Python restatements for the biome filter's checker tests.

Test infrastructure only. Two kinds of code live here:

* restatements of published algorithms, written from their specifications:
  xoroshiro128++ (Blackman & Vigna) with Minecraft's 128-bit seeding, and Ken
  Perlin's improved noise as Minecraft seeds it;
* readers of the reference's noise sources as TEXT (gpu_noise/
  MinecraftGPUNoise.cu and gpu_noise/spline_generated.cuh, read from the
  reference tree where it lies; nothing is compiled). The constants of the
  climate composition are taken from that text, and the generated terrain
  spline is transliterated statement by statement into Python over numpy
  float32/float64 scalars, keeping C's types: float literals and `const float`
  locals round to binary32, `(double)` casts and the double lerp to binary64.

`RefClimate` composes the climate of a quart column the way
MinecraftGPUNoise.cu:255-422 and :868-913 do, with every constant read from the
text, so a slip in the oracle's constants or steps (an octave count, a seed, an
amplitude, the double-Perlin factor, the shift, the quantization, the depth
terms) makes its values differ from this one.
"""
import math
import os
import re

import numpy as np

M64 = (1 << 64) - 1
REF_NOISE = "/root/reference/gpu_noise"


# -- published algorithms ----------------------------------------------------------

def rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def mix13(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


class Xoroshiro128PlusPlus:
    """Blackman & Vigna's xoroshiro128++, seeded as Minecraft's
    RandomSupport.upgradeSeedTo128bit (silver/golden ratio, Stafford mix 13)."""

    def __init__(self, seed, state=None):
        if state is not None:
            self.s0, self.s1 = state
            return
        lo = (seed ^ 0x6A09E667F3BCC909) & M64
        hi = (lo + 0x9E3779B97F4A7C15) & M64
        self.s0, self.s1 = mix13(lo), mix13(hi)

    @classmethod
    def from_state(cls, lo, hi):
        return cls(None, (lo & M64, hi & M64))

    def next_long(self):
        s0, s1 = self.s0, self.s1
        r = (rotl((s0 + s1) & M64, 17) + s0) & M64
        s1 ^= s0
        self.s0 = rotl(s0, 49) ^ s1 ^ ((s1 << 21) & M64)
        self.s1 = rotl(s1, 28)
        return r

    def next_int(self, n):
        while True:
            r = (self.next_long() & 0xFFFFFFFF) * n
            if (r & 0xFFFFFFFF) >= (-n % (1 << 32)) % n:
                return r >> 32

    def next_double(self):
        return (self.next_long() >> 11) * 2.0 ** -53


GRAD = [(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
        (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (0, -1, 1),
        (-1, 1, 0), (0, -1, -1)]


class ImprovedNoise:
    """Ken Perlin's improved noise as Minecraft seeds it (origin from three
    doubles, Fisher-Yates permutation)."""

    def __init__(self, rng):
        self.o = [rng.next_double() * 256.0 for _ in range(3)]
        p = list(range(256))
        for i in range(256):
            j = i + rng.next_int(256 - i)
            p[i], p[j] = p[j], p[i]
        self.p = p + [p[0]]

    @staticmethod
    def fade(t):
        return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)

    @staticmethod
    def lerp(t, a, b):
        return a + t * (b - a)

    def grad(self, h, x, y, z):
        g = GRAD[h & 15]
        terms = [v if s > 0 else -v for s, v in zip(g, (x, y, z)) if s]
        return terms[0] + terms[1]

    def sample(self, x, y, z):
        x, y, z = x + self.o[0], y + self.o[1], z + self.o[2]
        fx, fy, fz = math.floor(x), math.floor(y), math.floor(z)
        x, y, z = x - fx, y - fy, z - fz
        X, Y, Z = int(fx) & 255, int(fy) & 255, int(fz) & 255
        u, v, w = self.fade(x), self.fade(y), self.fade(z)
        p = self.p
        a, b = (p[X] + Y) & 255, (p[X + 1] + Y) & 255
        aa, ba, ab, bb = (p[a] + Z) & 255, (p[b] + Z) & 255, (p[a + 1] + Z) & 255, (p[b + 1] + Z) & 255
        L = self.lerp
        g = self.grad
        x0 = L(u, g(p[aa], x, y, z), g(p[ba], x - 1, y, z))
        x1 = L(u, g(p[ab], x, y - 1, z), g(p[bb], x - 1, y - 1, z))
        x2 = L(u, g(p[aa + 1], x, y, z - 1), g(p[ba + 1], x - 1, y, z - 1))
        x3 = L(u, g(p[ab + 1], x, y - 1, z - 1), g(p[bb + 1], x - 1, y - 1, z - 1))
        return L(w, L(v, x0, x1), L(v, x2, x3))


# -- the reference's noise source, read as text -------------------------------------

def ref_available():
    return os.path.exists(os.path.join(REF_NOISE, "MinecraftGPUNoise.cu"))


def _norm(s):
    return re.sub(r"\s+", " ", s)


def c_double_expr(expr):
    """Value of a C double constant expression made of decimal literals and
    '/' (e.g. `1.0/8.0`): one IEEE division, as C evaluates it."""
    expr = expr.strip()
    if not re.fullmatch(r"-?\d+\.\d*(\s*/\s*\d+\.\d*)?", expr):
        raise ValueError(f"unexpected constant expression {expr!r}")
    parts = [float(p) for p in expr.split("/")]
    return parts[0] / parts[1] if len(parts) == 2 else parts[0]


def function_body(text, name):
    """Text between the braces of the definition of `name` (brace-matched)."""
    m = re.search(r"\b" + re.escape(name) + r"\s*\([^;{]*\)\s*\{", text)
    if not m:
        raise ValueError(f"{name} not found")
    i, depth = m.end(), 1
    while depth:
        c = text[i]
        depth += (c == "{") - (c == "}")
        i += 1
    return text[m.end():i - 1]


class NoiseConstants:
    """Every constant of the climate composition, read from
    MinecraftGPUNoise.cu's text."""

    def __init__(self, path=None):
        text = open(path or os.path.join(REF_NOISE, "MinecraftGPUNoise.cu")).read()
        self.text = text

        def double_table(name, n):
            m = re.search(r"double " + name + r"\[" + str(n) + r"\]\s*=\s*\{([^}]*)\}", text)
            vals = [c_double_expr(v) for v in m.group(1).split(",") if v.strip()]
            assert len(vals) == n, name
            return vals

        self.lacuna = double_table("kLacunaInit", 17)      # :227-232
        self.persist = double_table("kPersistInit", 17)    # :234-240
        self.double_amp = double_table("kDoubleAmp", 17)   # :249-253
        self.amps = {}
        for m in re.finditer(r"double (k\w+Amp)\[(\d+)\]\s*=\s*\{([^}]*)\}", text):
            if m.group(1) != "kDoubleAmp":
                self.amps[m.group(1)] = [c_double_expr(v) for v in m.group(3).split(",")]
        m = re.search(r"uint64_t kMd5Octave\[17\]\[2\]\s*=\s*\{(.*?)\};", text, re.S)
        pairs = re.findall(r"\{0x([0-9a-fA-F]+)ULL,\s*0x([0-9a-fA-F]+)ULL\}", m.group(1))
        self.octave_seeds = [(int(a, 16), int(b, 16)) for a, b in pairs]
        assert len(self.octave_seeds) == 17
        # init_climate_state_mode (:369-405): per noise, the seed halves (normal
        # and large biomes), amplitude table, first octave and length
        body = _norm(function_body(text, "init_climate_state_mode"))
        hexn = r"0x([0-9a-fA-F]+)ULL"
        half = r"(?:\(large \? " + hexn + r" : " + hexn + r"\)|" + hexn + r")"
        omin = r"(?:large \? (-?\d+) : (-?\d+)|(-?\d+))"
        pat = (r"init_compact_dp\( st, C_(\w+), off, \{xlo \^ " + half + r", xhi \^ " + half +
               r"\}, (k\w+Amp), " + omin + r", (\d+)\);")
        self.noises = {}
        order = []
        for g in re.findall(pat, body):
            slot = g[0]
            lo_l, lo_n, lo_1, hi_l, hi_n, hi_1 = g[1:7]
            amp, om_l, om_n, om_1, length = g[7:12]
            lo = int(lo_n or lo_1, 16)
            hi = int(hi_n or hi_1, 16)
            self.noises[slot] = {
                "lo": lo, "hi": hi,
                "large_lo": int(lo_l, 16) if lo_l else lo,
                "large_hi": int(hi_l, 16) if hi_l else hi,
                "amps": self.amps[amp],
                "omin": int(om_n or om_1), "large_omin": int(om_l) if om_l else int(om_1),
                "len": int(length)}
            order.append(slot)
        assert order == ["SHIFT", "TEMPERATURE", "HUMIDITY", "CONTINENTALNESS", "EROSION",
                         "WEIRDNESS"], order
        self.init_order = order
        # the second octave set's input factor (sample_compact_dp, :413-422)
        sc = _norm(function_body(text, "sample_compact_dp"))
        m = re.search(r"const double f = (\d+\.\d*) / (\d+\.\d*);", sc)
        self.dp_factor = float(m.group(1)) / float(m.group(2))
        assert "v += sample_octaves(base + d.countA, d.countB, x*f, y*f, z*f); return v * d.amplitude;" in sc
        # biome_modern_at (:868-913): the shift, the quantization, the depth
        bm = _norm(function_body(text, "biome_modern_at"))
        m = re.search(r"const double px = x \+ sample_compact_dp\(st, C_SHIFT, x, 0\.0, z\) \* "
                      r"(\d+\.\d*); const double pz = z \+ sample_compact_dp\(st, C_SHIFT, z, x, "
                      r"0\.0\) \* (\d+\.\d*);", bm)
        self.shift = (float(m.group(1)), float(m.group(2)))
        quant = re.findall(r"np\[(\d)\] = \(int32_t\)\((\d+\.\d*)F \* (\w)\);", bm)
        self.quant_index = {v: int(i) for i, _, v in quant}   # climate letter -> np slot
        self.quant_scale = {float(s) for _, s, _ in quant}
        assert len(quant) == 6 and len(self.quant_scale) == 1
        m = re.search(r"const float vals\[4\] = \{ (\w), (\w), mcgpu_peaks_and_valleys\((\w)\), "
                      r"(\w) \};", bm)
        self.spline_args = m.groups()
        m = re.search(r"const float off_f = mcgpu_depth_spline\(vals\) \+ (\d+\.\d*)F;", bm)
        self.depth_offset = float(m.group(1))
        m = re.search(r"const double d0 = (\d+\.\d*) - \(\(double\)\(yi \* (\d+)\)\) / (\d+\.\d*) - "
                      r"(\d+\.\d*) / (\d+\.\d*) \+ \(double\)off_f;", bm)
        self.depth = tuple(float(x) for x in m.groups())
        order = re.findall(r"const float (\w) = \(float\)sample_compact_dp\( st, C_(\w+), px, 0\.0, pz\);", bm)
        self.climate_of = {v: slot for v, slot in order}
        pv = _norm(function_body(text, "mcgpu_peaks_and_valleys"))
        m = re.search(r"return -\(fabsf\(fabsf\(weirdness\) - (\d+\.\d*)F\) - (\d+\.\d*)F\) \* "
                      r"(\d+\.\d*)F;", pv)
        self.pv = tuple(m.groups())
        # the octave set (xoctave_init_profile, :255-282): lacunarity doubles,
        # persistence halves per octave, zero amplitudes are skipped
        xo = _norm(function_body(text, "xoctave_init_profile"))
        for s in ("double lacuna = kLacunaInit[-omin]; double persist = kPersistInit[len];",
                  "for (int i = 0; i < len; ++i, lacuna *= 2.0, persist *= 0.5) {",
                  "if (amp == 0.0) continue; const int md5i = 16 + omin + i;",
                  "xlo ^ kMd5Octave[md5i][0], xhi ^ kMd5Octave[md5i][1]",
                  "out[n].amplitude = amp * persist; out[n].lacunarity = lacuna;"):
            assert s in xo, s
        # the double-Perlin amplitude of a noise from its trimmed octave count
        ic = _norm(function_body(text, "init_compact_dp"))
        for s in ("while (trimmed > 0 && amplitudes[trimmed - 1] == 0.0) --trimmed;",
                  "while (leading < trimmed && amplitudes[leading] == 0.0) { ++leading; --trimmed; }",
                  "d.amplitude = kDoubleAmp[trimmed];"):
            assert s in ic, s


def f32(x):
    return np.float32(x)


def f64(x):
    return np.float64(x)


class RefClimate:
    """The climate of a world as MinecraftGPUNoise.cu composes it, from the
    constants its text holds (NoiseConstants) and the published Perlin and
    xoroshiro algorithms above."""

    def __init__(self, k, world_seed, large=False):
        self.k = k
        base = Xoroshiro128PlusPlus(world_seed & M64)
        xlo, xhi = base.next_long(), base.next_long()
        self.noise = {}
        for slot in k.init_order:
            d = k.noises[slot]
            xr = Xoroshiro128PlusPlus.from_state(xlo ^ (d["large_lo"] if large else d["lo"]),
                                                 xhi ^ (d["large_hi"] if large else d["hi"]))
            omin = d["large_omin"] if large else d["omin"]
            sets = [self._octaves(xr, d["amps"], omin, d["len"]) for _ in range(2)]
            amps = d["amps"][:d["len"]]
            trimmed = len(amps)
            while trimmed > 0 and amps[trimmed - 1] == 0.0:
                trimmed -= 1
            leading = 0
            while leading < trimmed and amps[leading] == 0.0:
                leading += 1
                trimmed -= 1
            self.noise[slot] = (sets, k.double_amp[trimmed])

    def _octaves(self, parent, amps, omin, length):
        lacuna, persist = self.k.lacuna[-omin], self.k.persist[length]
        xlo, xhi = parent.next_long(), parent.next_long()
        out = []
        for i in range(length):
            if amps[i] != 0.0:
                s = self.k.octave_seeds[16 + omin + i]
                out.append((ImprovedNoise(Xoroshiro128PlusPlus.from_state(xlo ^ s[0], xhi ^ s[1])),
                            amps[i] * persist, lacuna))
            lacuna *= 2.0
            persist *= 0.5
        return out

    @staticmethod
    def _octave_sum(octaves, x, y, z):
        v = 0.0
        for p, amp, lf in octaves:
            v += amp * p.sample(x * lf, y * lf, z * lf)
        return v

    def sample(self, slot, x, y, z):
        """sample_compact_dp: each octave set summed on its own, the second
        at the inputs scaled by the double-Perlin factor."""
        (a, b), amplitude = self.noise[slot]
        f = self.k.dp_factor
        v = self._octave_sum(a, x, y, z)
        v += self._octave_sum(b, x * f, y * f, z * f)
        return v * amplitude

    def shifted(self, x, z):
        sx, sz = self.k.shift
        return (x + self.sample("SHIFT", x, 0.0, z) * sx,
                z + self.sample("SHIFT", z, x, 0.0) * sz)

    def climate5(self, x, z):
        """Doubles t, h, c, e, w at the shifted point of (x, z)."""
        px, pz = self.shifted(x, z)
        return [self.sample(s, px, 0.0, pz)
                for s in ("TEMPERATURE", "HUMIDITY", "CONTINENTALNESS", "EROSION", "WEIRDNESS")]

    def column(self, spline, xi, zi):
        """The climate floats of the quart column (xi, zi) by letter (t, h,
        c, e, w) and the float depth offset of the terrain spline."""
        k = self.k
        px, pz = self.shifted(float(xi), float(zi))
        vals = {v: f32(self.sample(slot, px, 0.0, pz)) for v, slot in k.climate_of.items()}
        a, b, s = (f32(x) for x in k.pv)
        w = vals[k.spline_args[2]]
        pv = -(abs(abs(w) - a) - b) * s
        args = np.array([vals[k.spline_args[0]], vals[k.spline_args[1]], pv,
                         vals[k.spline_args[3]]], dtype=np.float32)
        return vals, spline(args) + f32(k.depth_offset)

    def params(self, column, yi):
        """The six quantized noise parameters of a column at quart y, in the
        tree's order."""
        k = self.k
        vals, off = column
        one, four, div, n83, d160 = k.depth
        d0 = one - float(yi * int(four)) / div - n83 / d160 + float(off)
        vals = dict(vals, d=f32(d0))
        scale = f32(next(iter(k.quant_scale)))
        np6 = [0] * 6
        for v, i in k.quant_index.items():
            np6[i] = int(scale * vals[v])
        return np6

    def climate6(self, spline, xi, yi, zi):
        col = self.column(spline, xi, zi)
        return self.params(col, yi), col[1]


class FilterSampling:
    """The candidate filter's sampling set, read from
    biome_filter_candidates_kernel (MinecraftGPUNoise.cu:978-1074): the disk,
    the quart columns of a chunk, the Deep Dark layers and the Mushroom Fields
    layer, and the two biome ids."""

    def __init__(self, text):
        body = _norm(function_body(text, "biome_filter_candidates_kernel"))
        m = re.search(r"const int dx = ci / (\d+) - (\d+); const int dz = ci % \1 - \2; "
                      r"if \(dx \* dx \+ dz \* dz <= (\d+) && slime_chunk_exact\(world_seed, "
                      r"ccx \+ dx, ccz \+ dz\)\)", body)
        self.side, self.half, self.r2 = int(m.group(1)), int(m.group(2)), int(m.group(3))
        self.quarts = int(re.search(r"constexpr int QUARTS_PER_CHUNK = (\d+);", body).group(1))
        self.layers = int(re.search(r"constexpr int DEEP_LAYERS = (\d+);", body).group(1))
        m = re.search(r"const int nx = quart_index / (\d+); const int nz = quart_index % \1; "
                      r"const int qx = \(ccx \+ dx\) \* (\d+) \+ nx; "
                      r"const int qz = \(ccz \+ dz\) \* \2 \+ nz; "
                      r"const int qy = mushroom \? (-?\d+) : \((-?\d+) \+ y_index\);", body)
        self.qside, self.qscale = int(m.group(1)), int(m.group(2))
        self.mushroom_y, self.deep_y0 = int(m.group(3)), int(m.group(4))
        m = re.search(r"if \(!mushroom && biome == (\d+)\) atomicOr\(&hit_mask, 1U\); "
                      r"else if \(mushroom && biome == (\d+)\) atomicOr\(&hit_mask, 2U\);", body)
        self.deep_dark, self.mushroom = int(m.group(1)), int(m.group(2))

    def mask(self, is_slime, biome_of, seed, cx, cz, flags):
        """The filter's mask of candidate (cx, cz): biome_of(qx, qy, qz)."""
        mask = 0
        for ci in range(self.side * self.side):
            dx, dz = ci // self.side - self.half, ci % self.side - self.half
            if dx * dx + dz * dz > self.r2 or not is_slime(seed, cx + dx, cz + dz):
                continue
            for q in range(self.quarts):
                qx = (cx + dx) * self.qscale + q // self.qside
                qz = (cz + dz) * self.qscale + q % self.qside
                if flags & 1 and any(biome_of(qx, self.deep_y0 + y, qz) == self.deep_dark
                                     for y in range(self.layers)):
                    mask |= 1
                if flags & 2 and biome_of(qx, self.mushroom_y, qz) == self.mushroom:
                    mask |= 2
        return mask


# -- the terrain spline, transliterated from spline_generated.cuh --------------------

_FLOAT_LIT = re.compile(r"(\d+\.\d*(?:e[+-]?\d+)?)f\b")


def _c_expr(e):
    e = _FLOAT_LIT.sub(r"F32('\1')", e)
    e = re.sub(r"\(double\)(\w+)", r"F64(\1)", e)
    e = e.replace("(float)(", "F32(")
    e = re.sub(r"\bspline_(\d+)\(v\)", r"S[\1](v)", e)
    e = e.replace("spline_lerp_d(", "LERP(")
    return e


def spline_functions(path=None):
    """{n: python function} for every spline_<n>(const float* v) of the
    reference's generated spline source, and the function mcgpu_depth_spline
    calls. Each C statement becomes one Python statement."""
    text = open(path or os.path.join(REF_NOISE, "spline_generated.cuh")).read()
    lerp = _norm(function_body(text, "spline_lerp_d"))
    assert lerp.strip() == "return a + t * (b - a);", lerp
    env = {"F32": np.float32, "F64": np.float64, "S": {},
           "LERP": lambda t, a, b: a + t * (b - a)}
    for m in re.finditer(r"__device__ __forceinline__ float spline_(\d+)\(const float\* v\)\s*\{",
                         text):
        n = int(m.group(1))
        body = function_body(text[m.start():], f"spline_{n}")
        lines, ind = [f"def spline_{n}(v):"], 1
        for raw in body.split("\n"):
            s = raw.strip()
            if not s:
                continue
            if s == "}":
                ind -= 1
                continue
            mm = re.fullmatch(r"if \((\w+) <= (.*)\) \{", s)
            if mm:
                lines.append("    " * ind + f"if {mm.group(1)} <= {_c_expr(mm.group(2))}:")
                ind += 1
                continue
            mm = re.fullmatch(r"const (float|double) (\w+) = (.*);", s)
            if mm:
                cast = "F32" if mm.group(1) == "float" else "F64"
                lines.append("    " * ind + f"{mm.group(2)} = {cast}({_c_expr(mm.group(3))})")
                continue
            mm = re.fullmatch(r"return (.*);", s)
            if mm:
                lines.append("    " * ind + f"return F32({_c_expr(mm.group(1))})")
                continue
            raise ValueError(f"spline_{n}: unexpected statement {s!r}")
        exec(compile("\n".join(lines), f"spline_{n}", "exec"), env)
        env["S"][n] = env[f"spline_{n}"]
    top = _norm(function_body(text, "mcgpu_depth_spline"))
    m = re.fullmatch(r" ?return spline_(\d+)\(v\); ?", top)
    return env["S"], env["S"][int(m.group(1))]
