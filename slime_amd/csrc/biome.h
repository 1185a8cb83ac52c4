// Minecraft's overworld multi-noise biomes (the 1.18+ climate sampler with
// the parameter trees of 1.18 to 26.2), as host and device code.
//
// A biome at quart coordinates (x, y, z) (blocks / 4) is the leaf of the
// parameter search tree nearest to the point
//   (temperature, humidity, continentalness, erosion, depth, weirdness),
// each a quantized climate value: five double-precision "double Perlin"
// noises sampled at the shifted point (px, pz) and the depth, a terrain
// spline of continentalness, erosion and peaks-and-valleys plus a y term.
// Reference: MinecraftGPUNoise.cu:15-422 (xoroshiro, Perlin, octaves, climate
// state), :657-913 (peaks and valleys, tree search, biome_modern_at) and
// spline_generated.cuh (the terrain spline, here evaluated from data).
//
// Only the five climate noises' y-independence is used beyond the reference:
// they are sampled at y = 0, so one climate evaluation serves every y layer
// of a quart column; the depth alone varies with y.
#pragma once

#include <math.h>
#include <stdint.h>

#include "slime_math.h"

namespace slime {
namespace biome {

struct BiomeNoiseDef {
    uint64_t lo, hi, large_lo, large_hi;  // MD5 halves of the noise name
    int32_t omin, large_omin, len;        // first octave, octave count
    double amp[9];
};
struct BiomeSpline {
    int32_t coord, first, count;
};
struct BiomeSplinePoint {
    uint32_t loc, der, value;  // binary32 bits
    int32_t child;             // sub-spline giving the value, or -1
};
struct BiomeTreeDef {
    int32_t version;        // first version code (e.g. 1206 = 1.20.6) using the tree
    int32_t family;         // node layout: 0 (1.18, 1.19.2) or 1 (1.19.4 on)
    const uint64_t* nodes;
    int32_t len;
    const int32_t* params;  // (lo, hi) pairs
    int32_t nparams;
};

// Tables that device code reads live in constant memory on the device.
#if defined(__HIP_DEVICE_COMPILE__)
#define BIOME_TABLE static __constant__ const
#else
#define BIOME_TABLE static const
#endif
#include "biome_data.inc"
#undef BIOME_TABLE

constexpr int kTreeCount = (int)(sizeof(kTreeDefs) / sizeof(kTreeDefs[0]));
constexpr int kTreeMaxParams = BIOME_TREE_MAX_PARAMS;
constexpr int kClimateNoises = 6;      // shift, temperature, humidity, continentalness, erosion, weirdness
constexpr int kClimateOctaves = 46;    // octaves of all six double Perlin noises
constexpr int kDeepDark = 183;         // biome ids of the filter
constexpr int kMushroomFields = 14;
enum { kShift = 0, kTemperature = 1, kHumidity = 2, kContinentalness = 3, kErosion = 4, kWeirdness = 5 };

// ---- xoroshiro128++ with Minecraft's seeding ------------------------------------

struct Xoro {
    uint64_t lo, hi;
};

SLIME_HD uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

SLIME_HD uint64_t mix_stafford13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The 128-bit state of a 64-bit seed: lo = seed ^ silver ratio, hi = lo +
// golden ratio, both mixed.
SLIME_HD Xoro xoro_from_seed(uint64_t seed) {
    const uint64_t lo = seed ^ 0x6A09E667F3BCC909ull;
    const uint64_t hi = lo + 0x9E3779B97F4A7C15ull;
    return Xoro{mix_stafford13(lo), mix_stafford13(hi)};
}

SLIME_HD uint64_t xoro_next(Xoro& s) {
    const uint64_t lo = s.lo, hi0 = s.hi;
    const uint64_t out = rotl64(lo + hi0, 17) + lo;
    const uint64_t hi = hi0 ^ lo;
    s.lo = rotl64(lo, 49) ^ hi ^ (hi << 21);
    s.hi = rotl64(hi, 28);
    return out;
}

// Uniform in [0, n) by multiply-and-reject on the low 32 bits.
SLIME_HD uint32_t xoro_int(Xoro& s, uint32_t n) {
    uint64_t r = (xoro_next(s) & 0xFFFFFFFFull) * n;
    if ((uint32_t)r < n) {
        const uint32_t limit = (0u - n) % n;
        while ((uint32_t)r < limit) r = (xoro_next(s) & 0xFFFFFFFFull) * n;
    }
    return (uint32_t)(r >> 32);
}

SLIME_HD double xoro_double(Xoro& s) { return (double)(xoro_next(s) >> 11) * 0x1.0p-53; }

// ---- improved Perlin noise ---------------------------------------------------

struct Perlin {
    uint8_t perm[257];  // perm[256] == perm[0]
    uint8_t y0_cell;    // cell and fade of y = 0 (the common case)
    uint8_t pad[6];
    double ox, oy, oz;  // origin offsets
    double y0_frac, y0_fade;
    double amp, lac;    // octave amplitude and input scale
};

SLIME_HD double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
SLIME_HD double lerp(double t, double a, double b) { return a + t * (b - a); }

// Dot product with one of Perlin's 16 gradient vectors. Every vector has two
// non-zero unit components, so the product is (+-a) + (+-b) with a in {x, y}
// and b in {y, z}: selects and sign flips instead of a 16-way branch (the
// same IEEE result: x - y is x + (-y)).
SLIME_HD double grad(uint8_t h, double x, double y, double z) {
    const int k = h & 15;
    const double a = (0xAF00u >> k) & 1u ? y : x;        // hashes 8-11, 13, 15
    const double b = (0xAFF0u >> k) & 1u ? z : y;         // hashes 4-11, 13, 15
    const bool na = (0xEAAAu >> k) & 1u;                  // 1, 3, 5, 7, 9, 11, 13, 14, 15
    const bool nb = (0x8CCCu >> k) & 1u;                  // 2, 3, 6, 7, 10, 11, 15
    return (na ? -a : a) + (nb ? -b : b);
}

// Origins from three doubles, then a Fisher-Yates shuffle of 0..255.
SLIME_HD void perlin_init(Perlin& p, Xoro xr) {
    p.ox = xoro_double(xr) * 256.0;
    p.oy = xoro_double(xr) * 256.0;
    p.oz = xoro_double(xr) * 256.0;
    for (int i = 0; i < 256; ++i) p.perm[i] = (uint8_t)i;
    for (int i = 0; i < 256; ++i) {
        const int j = i + (int)xoro_int(xr, (uint32_t)(256 - i));
        const uint8_t t = p.perm[i];
        p.perm[i] = p.perm[j];
        p.perm[j] = t;
    }
    p.perm[256] = p.perm[0];
    const double cell = floor(p.oy);
    p.y0_cell = (uint8_t)(int)cell;
    p.y0_frac = p.oy - cell;
    p.y0_fade = fade(p.y0_frac);
}

template <typename PerlinRef>
SLIME_HD double perlin_sample(const PerlinRef& p, double x, double y, double z) {
    uint8_t cy;
    double ty;
    if (y == 0.0) {
        y = p.y0_frac;
        cy = p.y0_cell;
        ty = p.y0_fade;
    } else {
        y += p.oy;
        const double fy = floor(y);
        y -= fy;
        cy = (uint8_t)(int)fy;
        ty = fade(y);
    }
    x += p.ox;
    z += p.oz;
    const double fx = floor(x), fz = floor(z);
    x -= fx;
    z -= fz;
    const uint8_t cx = (uint8_t)(int)fx, cz = (uint8_t)(int)fz;
    const double tx = fade(x), tz = fade(z);
    const uint8_t a = (uint8_t)(p.perm[cx] + cy), b = (uint8_t)(p.perm[cx + 1] + cy);
    const uint8_t aa = (uint8_t)(p.perm[a] + cz), ba = (uint8_t)(p.perm[b] + cz);
    const uint8_t ab = (uint8_t)(p.perm[a + 1] + cz), bb = (uint8_t)(p.perm[b + 1] + cz);
    const double g000 = grad(p.perm[aa], x, y, z);
    const double g100 = grad(p.perm[ba], x - 1.0, y, z);
    const double g010 = grad(p.perm[ab], x, y - 1.0, z);
    const double g110 = grad(p.perm[bb], x - 1.0, y - 1.0, z);
    const double g001 = grad(p.perm[aa + 1], x, y, z - 1.0);
    const double g101 = grad(p.perm[ba + 1], x - 1.0, y, z - 1.0);
    const double g011 = grad(p.perm[ab + 1], x, y - 1.0, z - 1.0);
    const double g111 = grad(p.perm[bb + 1], x - 1.0, y - 1.0, z - 1.0);
    const double x00 = lerp(tx, g000, g100), x10 = lerp(tx, g010, g110);
    const double x01 = lerp(tx, g001, g101), x11 = lerp(tx, g011, g111);
    return lerp(tz, lerp(ty, x00, x10), lerp(ty, x01, x11));
}

// ---- climate: six double Perlin noises -------------------------------------------

struct ClimateNoise {
    int32_t first;     // first octave in Climate::oct
    int32_t count_a;   // octaves of the first set (the second follows)
    int32_t count_b;
    int32_t pad;
    double amplitude;
};

struct Climate {
    ClimateNoise noise[kClimateNoises];
    Perlin oct[kClimateOctaves];
};

// One octave set: the non-zero amplitudes of octaves omin .. omin+len-1, each
// seeded from the set's forked state and the octave's name.
inline int octaves_init(Perlin* out, Xoro& parent, const double* amp, int omin, int len) {
    const uint64_t lo = xoro_next(parent), hi = xoro_next(parent);
    double lac = ldexp(1.0, omin);
    double persist = (double)(1 << (len - 1)) / (double)((1 << len) - 1);
    int n = 0;
    for (int i = 0; i < len; ++i, lac *= 2.0, persist *= 0.5) {
        if (amp[i] == 0.0) continue;
        const int o = 16 + omin + i;  // octave name index, -16..0 -> 0..16
        perlin_init(out[n], Xoro{lo ^ kOctaveSeeds[o][0], hi ^ kOctaveSeeds[o][1]});
        out[n].amp = amp[i] * persist;
        out[n].lac = lac;
        ++n;
    }
    return n;
}

inline void climate_init(Climate& c, uint64_t world_seed, bool large) {
    Xoro base = xoro_from_seed(world_seed);
    const uint64_t lo = xoro_next(base), hi = xoro_next(base);
    int first = 0;
    for (int k = 0; k < kClimateNoises; ++k) {
        const BiomeNoiseDef& d = kBiomeNoises[k];
        Xoro xr{lo ^ (large ? d.large_lo : d.lo), hi ^ (large ? d.large_hi : d.hi)};
        const int omin = large ? d.large_omin : d.omin;
        ClimateNoise& cn = c.noise[k];
        cn.first = first;
        cn.count_a = octaves_init(c.oct + first, xr, d.amp, omin, d.len);
        cn.count_b = octaves_init(c.oct + first + cn.count_a, xr, d.amp, omin, d.len);
        cn.pad = 0;
        // amplitude 5/3 * s / (s + 1) over the span s of non-zero octaves
        int lo_i = 0, hi_i = d.len;
        while (hi_i > 0 && d.amp[hi_i - 1] == 0.0) --hi_i;
        while (lo_i < hi_i && d.amp[lo_i] == 0.0) ++lo_i;
        const int span = hi_i - lo_i;
        cn.amplitude = (double)(5 * span) / (double)(3 * (span + 1));
        first += cn.count_a + cn.count_b;
    }
}

template <typename ClimateRef>
SLIME_HD double octaves_sample(const ClimateRef& c, int first, int count, double x, double y,
                               double z) {
    double v = 0.0;
    for (int i = first; i < first + count; ++i) {
        const double s = c.oct[i].lac;
        v += c.oct[i].amp * perlin_sample(c.oct[i], x * s, y * s, z * s);
    }
    return v;
}

template <typename ClimateRef>
SLIME_HD double noise_sample(const ClimateRef& c, int k, double x, double y, double z) {
    const ClimateNoise& n = c.noise[k];
    const double f = 337.0 / 331.0;
    double v = octaves_sample(c, n.first, n.count_a, x, y, z);
    v += octaves_sample(c, n.first + n.count_a, n.count_b, x * f, y * f, z * f);
    return v * n.amplitude;
}

// The y-independent climate of quart column (x, z): quantized temperature,
// humidity, continentalness, erosion, weirdness, and the depth spline offset.
struct Column {
    int32_t t, h, c, e, w;
    float offset;
};

SLIME_HD float bits_float(uint32_t b) {
    union {
        uint32_t u;
        float f;
    } v;
    v.u = b;
    return v.f;
}

SLIME_HD float peaks_and_valleys(float w) {
    return -(fabsf(fabsf(w) - 0.6666667F) - 0.33333334F) * 3.0F;
}

// Cubic spline with linear ends (Minecraft's CubicSpline), values of inner
// points from sub-splines; depth-bounded (the tree is 3 deep).
template <int D>
SLIME_HD float spline_eval(int id, const float* v) {
    const BiomeSpline s = kSplines[id];
    const BiomeSplinePoint* pt = kSplinePoints + s.first;
    const float f = v[s.coord];
    auto value = [&](int i) -> float {
        if (pt[i].child < 0) return bits_float(pt[i].value);
        if constexpr (D < 3) return spline_eval<D + 1>(pt[i].child, v);
        return 0.0f;
    };
    const float loc0 = bits_float(pt[0].loc);
    if (f <= loc0) return value(0) + bits_float(pt[0].der) * (f - loc0);
    for (int i = 0; i + 1 < s.count; ++i) {
        const float h = bits_float(pt[i + 1].loc);
        if (f <= h) {
            const float g = bits_float(pt[i].loc);
            const float k = (f - g) / (h - g);
            const float l = bits_float(pt[i].der), m = bits_float(pt[i + 1].der);
            const float n = value(i), o = value(i + 1);
            const float p = l * (h - g) - (o - n);
            const float q = -m * (h - g) + (o - n);
            const double a0 = lerp((double)k, (double)n, (double)o);
            const float kk = k * (1.0f - k);
            const double b0 = lerp((double)k, (double)p, (double)q);
            return (float)(a0 + (double)kk * b0);
        }
    }
    const int last = s.count - 1;
    return value(last) + bits_float(pt[last].der) * (f - bits_float(pt[last].loc));
}

// The point the climate noises are sampled at: (x, z) moved by the shift
// noise (MinecraftGPUNoise.cu:891-894).
template <typename ClimateRef>
SLIME_HD void shifted_point(const ClimateRef& c, double x, double z, double& px, double& pz) {
    px = x + noise_sample(c, kShift, x, 0.0, z) * 4.0;
    pz = z + noise_sample(c, kShift, z, x, 0.0) * 4.0;
}

// Climate value as the tree compares it: (int32)(10000 * (float)v).
SLIME_HD int32_t quantize(double v) { return (int32_t)(10000.0F * (float)v); }

template <typename ClimateRef>
SLIME_HD Column column_at(const ClimateRef& c, int32_t xi, int32_t zi) {
    double px, pz;
    shifted_point(c, (double)xi, (double)zi, px, pz);
    const float cf = (float)noise_sample(c, kContinentalness, px, 0.0, pz);
    const float ef = (float)noise_sample(c, kErosion, px, 0.0, pz);
    const float wf = (float)noise_sample(c, kWeirdness, px, 0.0, pz);
    const float tf = (float)noise_sample(c, kTemperature, px, 0.0, pz);
    const float hf = (float)noise_sample(c, kHumidity, px, 0.0, pz);
    Column col;
    col.c = (int32_t)(10000.0F * cf);
    col.e = (int32_t)(10000.0F * ef);
    col.w = (int32_t)(10000.0F * wf);
    col.t = (int32_t)(10000.0F * tf);
    col.h = (int32_t)(10000.0F * hf);
    const float sv[4] = {cf, ef, peaks_and_valleys(wf), wf};
    col.offset = spline_eval<0>(0, sv) + 0.015F;
    return col;
}

// Quantized depth of quart layer yi.
SLIME_HD int32_t depth_at(float offset, int32_t yi) {
    const double d = 1.0 - ((double)(yi * 4)) / 128.0 - 83.0 / 160.0 + (double)offset;
    return (int32_t)(10000.0F * (float)d);
}

// ---- parameter trees ---------------------------------------------------------------
//
// Node (64 bits): bytes 0..5 index the parameter interval of each climate
// dimension; bits 48..63 are the first child's index (children are `step`
// apart) or, with the top byte 0xFF, a leaf whose byte 6 is the biome id.
// Two layouts (MinecraftGPUNoise.cu:662-790): family 0 (1.18, 1.19.2) has up
// to 10 children per node and steps 1111, 111, 11, 1; family 1 (1.19.4 on)
// up to 6 and steps 1555, 259, 43, 7, 1.

template <int F>
struct TreeShape;
template <>
struct TreeShape<0> {
    static constexpr int order = 10, depth = 4;
    static constexpr int step(int d) { return d == 0 ? 1111 : d == 1 ? 111 : d == 2 ? 11 : 1; }
};
template <>
struct TreeShape<1> {
    static constexpr int order = 6, depth = 5;
    static constexpr int step(int d) {
        return d == 0 ? 1555 : d == 1 ? 259 : d == 2 ? 43 : d == 3 ? 7 : 1;
    }
};

// The tree of a version code (MinecraftGPUNoise.cu:1785-1820), or -1 below
// 1.18 (code 1180).
inline int tree_for_version(int32_t version_code) {
    if (version_code < 1180) return -1;
    for (int i = 0; i < kTreeCount; ++i)
        if (version_code >= kTreeDefs[i].version) return i;
    return kTreeCount - 1;
}

// params: the intervals, flattened (lo, hi) pairs.
SLIME_HD uint64_t node_distance(const int32_t np[6], const uint64_t* nodes, const int32_t* params,
                                int idx) {
    const uint64_t node = nodes[idx];
    uint64_t ds = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const int pi = (int)((node >> (8 * i)) & 0xFFu);
        const int64_t lo = params[2 * pi], hi = params[2 * pi + 1], v = np[i];
        const int64_t d = v > hi ? v - hi : (v < lo ? lo - v : 0);
        ds += (uint64_t)(d * d);
    }
    return ds;
}

// Nearest leaf below idx (the search of MinecraftGPUNoise.cu:680-790): visit
// the children whose box is nearer than best_ds, in order, and keep a leaf
// only if strictly nearer. `best` is the leaf kept so far or -1: a child's
// search returns a different leaf only when it kept a strictly nearer one, so
// the result is the first leaf in visiting order at the least distance below
// the bound (boxes never lie farther than the leaves below them). The
// reference starts from node 0 with an unbounded distance and takes a
// returned leaf when its distance is below the bound; that differs from this
// only if its first descent (always the first child of every node) ended
// without a leaf, which no tree has (checked by the tests). `stop(leaf)`
// ends the search early once it returns true for a kept leaf (then `stopped`
// is set and that leaf returned).
struct NoStop {
    SLIME_HD bool operator()(int) const { return false; }
};

template <int F, int D, typename Stop = NoStop>
SLIME_HD int tree_nearest(const int32_t np[6], const uint64_t* nodes, const int32_t* params,
                          int len, int idx, int best, uint64_t best_ds, const Stop& stop,
                          bool& stopped) {
    if constexpr (D >= TreeShape<F>::depth) {
        return idx;
    } else {
        constexpr int step = TreeShape<F>::step(D);
        if (idx + step >= len)
            return tree_nearest<F, D + 1>(np, nodes, params, len, idx, best, best_ds, stop,
                                          stopped);
        int child = (int)((nodes[idx] >> 48) & 0xFFFFu);
        for (int k = 0; k < TreeShape<F>::order && child < len; ++k, child += step) {
            const uint64_t dc = node_distance(np, nodes, params, child);
            if (dc < best_ds) {
                const int leaf = tree_nearest<F, D + 1>(np, nodes, params, len, child, best,
                                                        best_ds, stop, stopped);
                if (stopped) return leaf;
                if (leaf != best) {  // a strictly nearer leaf below the child
                    best_ds = leaf == child ? dc : node_distance(np, nodes, params, leaf);
                    best = leaf;
                    if (stop(leaf)) {
                        stopped = true;
                        return leaf;
                    }
                }
            }
        }
        return best;
    }
}

// The leaf the search from the root finds: the first leaf in visiting order
// at the least distance. A known leaf `hint` (hint >= 0) only bounds the
// search (distance(hint) + 1, above the least distance); it is not taken until
// the search reaches it, so the result is the same leaf. Returns the leaf.
template <int F>
SLIME_HD int tree_leaf(const int32_t np[6], const uint64_t* nodes, const int32_t* params, int len,
                       int hint) {
    bool stopped = false;
    const uint64_t bound = hint < 0 ? ~0ull : node_distance(np, nodes, params, hint) + 1;
    return tree_nearest<F, 0>(np, nodes, params, len, 0, -1, bound, NoStop{}, stopped);
}

SLIME_HD int leaf_biome(const uint64_t* nodes, int leaf) {
    return (int)((nodes[leaf] >> 48) & 0xFFu);
}

// Biome at quart layer yi of a column; `hint` carries the last leaf found
// (a start bound for the next search) and is updated.
template <int F>
SLIME_HD int biome_at_column(const Column& col, int32_t yi, const uint64_t* nodes,
                             const int32_t* params, int len, int& hint) {
    const int32_t np[6] = {col.t, col.h, col.c, col.e, depth_at(col.offset, yi), col.w};
    hint = tree_leaf<F>(np, nodes, params, len, hint);
    return leaf_biome(nodes, hint);
}

}  // namespace biome
}  // namespace slime
