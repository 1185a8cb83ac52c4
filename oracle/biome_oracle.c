/* CPU restatement of the biome filter of the slime-cluster search.
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * CPU leg call it as the checker; the product never links it.
 *
 * Follows the reference's gpu_noise/MinecraftGPUNoise.cu, which restates
 * Minecraft's multi-noise overworld biome source after cubiomes:
 *   xoroshiro128++ and its seeding        :15-70
 *   improved Perlin noise                 :72-190
 *   octave sets, double Perlin, climate   :192-422
 *   peaks and valleys, tree searches      :657-809
 *   version -> tree                       :1785-1820
 *   depth and biome at a quart            :868-913
 *   the candidate filter                  :945-1074
 * and evaluates the terrain spline (spline_generated.cuh) from the point data
 * of slime_amd/csrc/biome_data.inc. Compiled with -ffp-contract=off so every
 * double and float operation rounds as written.
 *
 * Pinning (tests/test_biome_oracle.py, tests/test_biome_reference_text.py):
 * the noise seeds are MD5 digests of the noise names and equal the
 * reference's text; xoroshiro and Perlin against independent Python
 * restatements of the published algorithms; every constant of the climate
 * composition (octave seeds, amplitudes, first octaves, octave counts,
 * lacunarity, persistence, double-Perlin amplitude and factor, shift,
 * quantization, depth terms) is read from MinecraftGPUNoise.cu's text, and
 * the climate doubles, quantized parameters and spline offsets of this file
 * equal a Python composition built from those constants bit for bit; the
 * spline equals spline_generated.cuh's text transliterated into numpy
 * float32/float64; the filter's sampling set equals the reference kernel's;
 * the tree table equals the reference's table text; the pruned tree search
 * equals an exhaustive nearest-leaf search, and the search started from a
 * known leaf equals the search from the root. Biome ids against the
 * reference's CUDA build (nvcc contracts multiply-adds into FMAs by default)
 * or cubiomes: parity unpinned (neither runs here). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t lo, hi, large_lo, large_hi;
    int32_t omin, large_omin, len;
    double amp[9];
} BiomeNoiseDef;
typedef struct {
    int32_t coord, first, count;
} BiomeSpline;
typedef struct {
    uint32_t loc, der, value;
    int32_t child;
} BiomeSplinePoint;
typedef struct {
    int32_t version, family;
    const uint64_t* nodes;
    int32_t len;
    const int32_t* params;
    int32_t nparams;
} BiomeTreeDef;
#define BIOME_TABLE static const
#include "../slime_amd/csrc/biome_data.inc"
#undef BIOME_TABLE

/* A parameter tree: nodes, intervals and node layout (family). */
typedef struct {
    const uint64_t* nodes;
    const int32_t (*params)[2];
    int len, family;
} ob_tree;

#define OB_TREE(sfx, fam) \
    { kTreeNodes##sfx, kTreeParams##sfx, (int)(sizeof(kTreeNodes##sfx) / sizeof(uint64_t)), fam }
/* The version selection of MinecraftGPUNoise.cu:1785-1820, trees by name. */
static const ob_tree kObTrees[7] = {OB_TREE(262, 1), OB_TREE(215, 1), OB_TREE(21wd, 1),
                                    OB_TREE(20, 1),  OB_TREE(19, 1),  OB_TREE(192, 0),
                                    OB_TREE(18, 0)};
static const int kObTreeFirstVersion[7] = {2602, 1215, 1214, 1206, 1194, 1192, 1180};

int ob_tree_for_version(int version) {
    if (version < 1180) return -1;
    for (int i = 0; i < 7; ++i)
        if (version >= kObTreeFirstVersion[i]) return i;
    return 6;
}

/* ---- xoroshiro128++ (MinecraftGPUNoise.cu:15-70) ---- */

typedef struct {
    uint64_t lo, hi;
} ob_xoro;

static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

static uint64_t stafford13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

ob_xoro ob_xoro_seed(uint64_t seed) {
    ob_xoro s;
    uint64_t lo = seed ^ 0x6A09E667F3BCC909ull;
    uint64_t hi = lo + 0x9E3779B97F4A7C15ull;
    s.lo = stafford13(lo);
    s.hi = stafford13(hi);
    return s;
}

uint64_t ob_xoro_next(ob_xoro* s) {
    uint64_t lo = s->lo, hi = s->hi;
    uint64_t r = rotl(lo + hi, 17) + lo;
    hi ^= lo;
    s->lo = rotl(lo, 49) ^ hi ^ (hi << 21);
    s->hi = rotl(hi, 28);
    return r;
}

uint32_t ob_xoro_int(ob_xoro* s, uint32_t n) {
    uint64_t r = (ob_xoro_next(s) & 0xFFFFFFFFull) * n;
    if ((uint32_t)r < n) {
        uint32_t t = (uint32_t)(-n) % n;
        while ((uint32_t)r < t) r = (ob_xoro_next(s) & 0xFFFFFFFFull) * n;
    }
    return (uint32_t)(r >> 32);
}

double ob_xoro_double(ob_xoro* s) { return (double)(ob_xoro_next(s) >> 11) * 1.1102230246251565E-16; }

/* First `count` outputs of seed's generator (for the tests). */
void ob_xoro_stream(uint64_t seed, int count, uint64_t* out) {
    ob_xoro s = ob_xoro_seed(seed);
    for (int i = 0; i < count; ++i) out[i] = ob_xoro_next(&s);
}

/* ---- improved Perlin noise (:72-190) ---- */

typedef struct {
    uint8_t d[257];
    double a, b, c;
    double amp, lac;
} ob_perlin;

static double smooth(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
static double mix(double t, double a, double b) { return a + t * (b - a); }

static double gradient(int h, double x, double y, double z) {
    /* Perlin's 16 gradient directions */
    static const int g[16][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0},
                                 {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
                                 {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1},
                                 {1, 1, 0}, {0, -1, 1}, {-1, 1, 0}, {0, -1, -1}};
    const int* v = g[h & 15];
    /* the two non-zero terms, in x, y, z order, exactly as two signed adds */
    double t[2];
    int n = 0;
    if (v[0]) t[n++] = v[0] > 0 ? x : -x;
    if (v[1]) t[n++] = v[1] > 0 ? y : -y;
    if (v[2]) t[n++] = v[2] > 0 ? z : -z;
    return t[0] + t[1];
}

void ob_perlin_init(ob_perlin* p, ob_xoro xr) {
    p->a = ob_xoro_double(&xr) * 256.0;
    p->b = ob_xoro_double(&xr) * 256.0;
    p->c = ob_xoro_double(&xr) * 256.0;
    for (int i = 0; i < 256; ++i) p->d[i] = (uint8_t)i;
    for (int i = 0; i < 256; ++i) {
        int j = i + (int)ob_xoro_int(&xr, (uint32_t)(256 - i));
        uint8_t t = p->d[i];
        p->d[i] = p->d[j];
        p->d[j] = t;
    }
    p->d[256] = p->d[0];
}

double ob_perlin_sample(const ob_perlin* p, double x, double y, double z) {
    x += p->a;
    y += p->b;
    z += p->c;
    double ix = floor(x), iy = floor(y), iz = floor(z);
    x -= ix;
    y -= iy;
    z -= iz;
    uint8_t h1 = (uint8_t)(int)ix, h2 = (uint8_t)(int)iy, h3 = (uint8_t)(int)iz;
    double t1 = smooth(x), t2 = smooth(y), t3 = smooth(z);
    const uint8_t* d = p->d;
    uint8_t a1 = (uint8_t)(d[h1] + h2), b1 = (uint8_t)(d[h1 + 1] + h2);
    uint8_t a2 = (uint8_t)(d[a1] + h3), b2 = (uint8_t)(d[b1] + h3);
    uint8_t a3 = (uint8_t)(d[a1 + 1] + h3), b3 = (uint8_t)(d[b1 + 1] + h3);
    double l1 = gradient(d[a2], x, y, z);
    double l2 = gradient(d[b2], x - 1.0, y, z);
    double l3 = gradient(d[a3], x, y - 1.0, z);
    double l4 = gradient(d[b3], x - 1.0, y - 1.0, z);
    double l5 = gradient(d[a2 + 1], x, y, z - 1.0);
    double l6 = gradient(d[b2 + 1], x - 1.0, y, z - 1.0);
    double l7 = gradient(d[a3 + 1], x, y - 1.0, z - 1.0);
    double l8 = gradient(d[b3 + 1], x - 1.0, y - 1.0, z - 1.0);
    l1 = mix(t1, l1, l2);
    l3 = mix(t1, l3, l4);
    l5 = mix(t1, l5, l6);
    l7 = mix(t1, l7, l8);
    l1 = mix(t2, l1, l3);
    l5 = mix(t2, l5, l7);
    return mix(t3, l1, l5);
}

/* One octave of seed's generator, sampled (for the tests). */
double ob_perlin_sample_seeded(uint64_t seed, double x, double y, double z) {
    ob_perlin p;
    ob_perlin_init(&p, ob_xoro_seed(seed));
    return ob_perlin_sample(&p, x, y, z);
}

/* ---- climate (:192-422) ---- */

typedef struct {
    int first, na, nb;
    double amplitude;
} ob_noise;

typedef struct {
    ob_noise noise[6];
    ob_perlin oct[46];
} ob_climate;

static int octave_set(ob_perlin* out, ob_xoro* parent, const double* amp, int omin, int len) {
    uint64_t lo = ob_xoro_next(parent), hi = ob_xoro_next(parent);
    double lac = ldexp(1.0, omin);
    double persist = (double)(1 << (len - 1)) / (double)((1 << len) - 1);
    int n = 0;
    for (int i = 0; i < len; ++i, lac *= 2.0, persist *= 0.5) {
        if (amp[i] == 0.0) continue;
        ob_xoro xr;
        xr.lo = lo ^ kOctaveSeeds[16 + omin + i][0];
        xr.hi = hi ^ kOctaveSeeds[16 + omin + i][1];
        ob_perlin_init(&out[n], xr);
        out[n].amp = amp[i] * persist;
        out[n].lac = lac;
        ++n;
    }
    return n;
}

ob_climate* ob_climate_new(uint64_t world_seed, int large) {
    ob_climate* c = (ob_climate*)calloc(1, sizeof(ob_climate));
    ob_xoro base = ob_xoro_seed(world_seed);
    uint64_t lo = ob_xoro_next(&base), hi = ob_xoro_next(&base);
    int first = 0;
    for (int k = 0; k < 6; ++k) {
        const BiomeNoiseDef* d = &kBiomeNoises[k];
        ob_xoro xr;
        xr.lo = lo ^ (large ? d->large_lo : d->lo);
        xr.hi = hi ^ (large ? d->large_hi : d->hi);
        int omin = large ? d->large_omin : d->omin;
        c->noise[k].first = first;
        c->noise[k].na = octave_set(c->oct + first, &xr, d->amp, omin, d->len);
        c->noise[k].nb = octave_set(c->oct + first + c->noise[k].na, &xr, d->amp, omin, d->len);
        int top = d->len, low = 0;
        while (top > 0 && d->amp[top - 1] == 0.0) --top;
        while (low < top && d->amp[low] == 0.0) ++low;
        c->noise[k].amplitude = (double)(5 * (top - low)) / (double)(3 * (top - low + 1));
        first += c->noise[k].na + c->noise[k].nb;
    }
    return c;
}

void ob_climate_free(ob_climate* c) { free(c); }

static double octaves(const ob_perlin* o, int n, double x, double y, double z) {
    double v = 0.0;
    for (int i = 0; i < n; ++i) {
        double l = o[i].lac;
        v += o[i].amp * ob_perlin_sample(&o[i], x * l, y * l, z * l);
    }
    return v;
}

static double noise(const ob_climate* c, int k, double x, double y, double z) {
    const ob_noise* n = &c->noise[k];
    const double f = 337.0 / 331.0;
    double v = octaves(c->oct + n->first, n->na, x, y, z);
    v += octaves(c->oct + n->first + n->na, n->nb, x * f, y * f, z * f);
    return v * n->amplitude;
}

/* ---- terrain spline (spline_generated.cuh, from its point data) ---- */

static float fbits(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static float spline_value(int id, const float* v);

static float point_value(const BiomeSplinePoint* p, const float* v) {
    return p->child < 0 ? fbits(p->value) : spline_value(p->child, v);
}

static float spline_value(int id, const float* v) {
    const BiomeSpline* s = &kSplines[id];
    const BiomeSplinePoint* pt = kSplinePoints + s->first;
    float f = v[s->coord];
    if (f <= fbits(pt[0].loc)) return point_value(&pt[0], v) + fbits(pt[0].der) * (f - fbits(pt[0].loc));
    for (int i = 0; i + 1 < s->count; ++i) {
        float h = fbits(pt[i + 1].loc);
        if (f > h) continue;
        float g = fbits(pt[i].loc);
        float k = (f - g) / (h - g);
        float l = fbits(pt[i].der), m = fbits(pt[i + 1].der);
        float n = point_value(&pt[i], v), o = point_value(&pt[i + 1], v);
        float p = l * (h - g) - (o - n);
        float q = -m * (h - g) + (o - n);
        double a0 = mix((double)k, (double)n, (double)o);
        float kk = k * (1.0f - k);
        double b0 = mix((double)k, (double)p, (double)q);
        return (float)(a0 + (double)kk * b0);
    }
    const BiomeSplinePoint* last = &pt[s->count - 1];
    return point_value(last, v) + fbits(last->der) * (f - fbits(last->loc));
}

float ob_spline(const float* v4) { return spline_value(0, v4); }

/* ---- parameter trees (:662-809) ---- */

uint64_t ob_node_distance(const int32_t* np, int idx, int tree) {
    const ob_tree* t = &kObTrees[tree];
    uint64_t node = t->nodes[idx], ds = 0;
    for (int i = 0; i < 6; ++i) {
        int pi = (int)((node >> (8 * i)) & 0xFF);
        int64_t lo = t->params[pi][0], hi = t->params[pi][1], v = np[i], d = 0;
        if (v > hi) d = v - hi;
        else if (v < lo) d = lo - v;
        ds += (uint64_t)(d * d);
    }
    return ds;
}

/* Family 1: up to 6 children, steps 1555, 259, 43, 7, 1 (btree20_search);
 * family 0: up to 10 children, steps 1111, 111, 11, 1 (btree18_search). */
static const int kSteps1[5] = {1555, 259, 43, 7, 1};
static const int kSteps0[4] = {1111, 111, 11, 1};

static int nearest(const int32_t* np, int tree, int depth, int idx, int best, uint64_t* best_ds) {
    const ob_tree* t = &kObTrees[tree];
    const int depths = t->family ? 5 : 4, order = t->family ? 6 : 10;
    if (depth >= depths) return idx;
    int step = t->family ? kSteps1[depth] : kSteps0[depth];
    if (idx + step >= t->len) return nearest(np, tree, depth + 1, idx, best, best_ds);
    int child = (int)((t->nodes[idx] >> 48) & 0xFFFF);
    uint64_t ds = *best_ds;
    for (int k = 0; k < order && child < t->len; ++k, child += step) {
        uint64_t dc = ob_node_distance(np, child, tree);
        if (dc < ds) {
            uint64_t sub = ds;
            int leaf = nearest(np, tree, depth + 1, child, best, &sub);
            uint64_t dl = leaf == child ? dc : ob_node_distance(np, leaf, tree);
            if (dl < ds) {
                ds = dl;
                best = leaf;
            }
        }
    }
    *best_ds = ds;
    return best;
}

/* The pruned search's leaf, and an exhaustive one: the first leaf (in table
 * order) at the least distance. */
int ob_tree_leaf(const int32_t* np, int tree) {
    uint64_t ds = ~0ull;
    return nearest(np, tree, 0, 0, 0, &ds);
}

/* The search bounded by a known leaf (what the kernels do from one y layer
 * to the next): no leaf kept at the start, the bound distance(hint) + 1, and
 * a child's leaf kept only when it differs from the current one (a strictly
 * nearer leaf was found below it). */
static int bounded(const int32_t* np, int tree, int depth, int idx, int best, uint64_t* best_ds) {
    const ob_tree* t = &kObTrees[tree];
    const int depths = t->family ? 5 : 4, order = t->family ? 6 : 10;
    if (depth >= depths) return idx;
    int step = t->family ? kSteps1[depth] : kSteps0[depth];
    if (idx + step >= t->len) return bounded(np, tree, depth + 1, idx, best, best_ds);
    int child = (int)((t->nodes[idx] >> 48) & 0xFFFF);
    for (int k = 0; k < order && child < t->len; ++k, child += step) {
        uint64_t dc = ob_node_distance(np, child, tree);
        if (dc < *best_ds) {
            uint64_t sub = *best_ds;
            int leaf = bounded(np, tree, depth + 1, child, best, &sub);
            if (leaf != best) {
                *best_ds = leaf == child ? dc : ob_node_distance(np, leaf, tree);
                best = leaf;
            }
        }
    }
    return best;
}

int ob_tree_leaf_from(const int32_t* np, int hint, int tree) {
    uint64_t ds = ob_node_distance(np, hint, tree) + 1;
    return bounded(np, tree, 0, 0, -1, &ds);
}

/* The first descent of the reference's search (first child at every level)
 * ends on a leaf: returns that node, so a test can check it is a leaf. */
int ob_tree_first_descent(int tree) {
    const ob_tree* t = &kObTrees[tree];
    const int depths = t->family ? 5 : 4;
    int idx = 0;
    for (int depth = 0; depth < depths; ++depth) {
        int step = t->family ? kSteps1[depth] : kSteps0[depth];
        if (idx + step >= t->len) continue;
        int child = (int)((t->nodes[idx] >> 48) & 0xFFFF);
        if (child >= t->len) return -1;
        idx = child;
    }
    return idx;
}

int ob_tree_leaf_exhaustive(const int32_t* np, int tree) {
    const ob_tree* t = &kObTrees[tree];
    int best = -1;
    uint64_t bd = ~0ull;
    for (int i = 0; i < t->len; ++i) {
        if ((t->nodes[i] >> 56) != 0xFF) continue;
        uint64_t d = ob_node_distance(np, i, tree);
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

int ob_leaf_biome(int leaf, int tree) { return (int)((kObTrees[tree].nodes[leaf] >> 48) & 0xFF); }
int ob_tree_len(int tree) { return kObTrees[tree].len; }
int ob_tree_family(int tree) { return kObTrees[tree].family; }
/* Interval i of a tree (lo, hi) and node i, for the table checks. */
void ob_tree_param(int tree, int i, int32_t* lo_hi) {
    lo_hi[0] = kObTrees[tree].params[i][0];
    lo_hi[1] = kObTrees[tree].params[i][1];
}
uint64_t ob_tree_node(int tree, int i) { return kObTrees[tree].nodes[i]; }

/* ---- climate of a quart column, biome at a quart (:868-913) ---- */

static float pv_of(float w) { return -(fabsf(fabsf(w) - 0.6666667F) - 0.33333334F) * 3.0F; }

/* out6: quantized t, h, c, e, w and the bits of the float depth offset. */
void ob_column(const ob_climate* c, int32_t xi, int32_t zi, int32_t* out6) {
    double x = (double)xi, z = (double)zi;
    double px = x + noise(c, 0, x, 0.0, z) * 4.0;
    double pz = z + noise(c, 0, z, x, 0.0) * 4.0;
    float cf = (float)noise(c, 3, px, 0.0, pz);
    float ef = (float)noise(c, 4, px, 0.0, pz);
    float wf = (float)noise(c, 5, px, 0.0, pz);
    float tf = (float)noise(c, 1, px, 0.0, pz);
    float hf = (float)noise(c, 2, px, 0.0, pz);
    float v[4] = {cf, ef, pv_of(wf), wf};
    float off = ob_spline(v) + 0.015F;
    out6[0] = (int32_t)(10000.0F * tf);
    out6[1] = (int32_t)(10000.0F * hf);
    out6[2] = (int32_t)(10000.0F * cf);
    out6[3] = (int32_t)(10000.0F * ef);
    out6[4] = (int32_t)(10000.0F * wf);
    memcpy(&out6[5], &off, 4);
}

static int32_t depth_of(const int32_t* col, int32_t yi) {
    float off;
    memcpy(&off, &col[5], 4);
    double d0 = 1.0 - ((double)(yi * 4)) / 128.0 - 83.0 / 160.0 + (double)off;
    return (int32_t)(10000.0F * (float)d0);
}

static int biome_from_column(const int32_t* col, int32_t yi, int tree) {
    int32_t np[6] = {col[0], col[1], col[2], col[3], depth_of(col, yi), col[4]};
    return ob_leaf_biome(ob_tree_leaf(np, tree), tree);
}

int ob_biome_at(const ob_climate* c, int32_t x, int32_t y, int32_t z, int tree) {
    int32_t col[6];
    ob_column(c, x, z, col);
    return biome_from_column(col, y, tree);
}

/* The point-query layouts (:532-655, :811-866). out5: t, h, c, e, w as
 * doubles at the shifted point of (x, z); out6: the quantized t, h, c, e,
 * depth at quart y, w. */
void ob_climate5(const ob_climate* c, double x, double z, double* out5) {
    double px = x + noise(c, 0, x, 0.0, z) * 4.0;
    double pz = z + noise(c, 0, z, x, 0.0) * 4.0;
    out5[0] = noise(c, 1, px, 0.0, pz);
    out5[1] = noise(c, 2, px, 0.0, pz);
    out5[2] = noise(c, 3, px, 0.0, pz);
    out5[3] = noise(c, 4, px, 0.0, pz);
    out5[4] = noise(c, 5, px, 0.0, pz);
}

void ob_climate6(const ob_climate* c, int32_t x, int32_t y, int32_t z, int32_t* out6) {
    int32_t col[6];
    ob_column(c, x, z, col);
    out6[0] = col[0];
    out6[1] = col[1];
    out6[2] = col[2];
    out6[3] = col[3];
    out6[4] = depth_of(col, y);
    out6[5] = col[4];
}

/* The continentalness noise alone at (x, 0, z), unshifted (:491-517). */
double ob_continentalness(const ob_climate* c, double x, double z) {
    return noise(c, 3, x, 0.0, z);
}

/* ---- the candidate filter (:945-1074) ---- */

static int slime(int64_t seed, int32_t x, int32_t z) {
    uint32_t ux = (uint32_t)x, uz = (uint32_t)z;
    uint64_t s = (uint64_t)seed + (uint64_t)(int64_t)(int32_t)(ux * ux * 0x4c1906u) +
                 (uint64_t)(int64_t)(int32_t)(ux * 0x5ac0dbu) +
                 (uint64_t)(int64_t)(int32_t)(uz * uz) * 0x4307a7ull +
                 (uint64_t)(int64_t)(int32_t)(uz * 0x5f24fu);
    uint64_t r = (s ^ 0x5E434E432ull) & 0xFFFFFFFFFFFFull;
    for (;;) {
        r = (r * 0x5DEECE66Dull + 0xBull) & 0xFFFFFFFFFFFFull;
        uint32_t bits = (uint32_t)(r >> 17), val = bits % 10u;
        if (bits - val + 9u <= 0x7FFFFFFFu) return val == 0u;
    }
}

int ob_filter(const ob_climate* c, int64_t world_seed, int32_t cx, int32_t cz, int flags,
              int tree) {
    int mask = 0;
    for (int ci = 0; ci < 17 * 17; ++ci) {
        int dx = ci / 17 - 8, dz = ci % 17 - 8;
        if (dx * dx + dz * dz > 68) continue;
        int32_t x = (int32_t)((uint32_t)cx + (uint32_t)dx), z = (int32_t)((uint32_t)cz + (uint32_t)dz);
        if (!slime(world_seed, x, z)) continue;
        for (int q = 0; q < 16; ++q) {
            if ((mask & flags) == flags) return mask;
            int32_t col[6];
            ob_column(c, (int32_t)((uint32_t)x * 4u + (uint32_t)(q / 4)),
                      (int32_t)((uint32_t)z * 4u + (uint32_t)(q % 4)), col);
            if ((flags & 2) && biome_from_column(col, 16, tree) == 14) mask |= 2;
            if (flags & 1)
                for (int y = -16; y <= 0; ++y)
                    if (biome_from_column(col, y, tree) == 183) {
                        mask |= 1;
                        break;
                    }
        }
    }
    return mask;
}

/* The filter over a level range, an extension of the reference's sampling
 * set (the wider sweep of y -64..320 step 4): flags bit 0 Deep Dark, bit 1
 * Mushroom Fields, bit 2 every chunk of the disk instead of its slime chunks.
 * Each requested biome is looked for at block heights y_min, y_min + y_step,
 * ... <= y_max, sampled at quart y floor(y / 4), in the 16 quart columns of
 * each sampled chunk. With levels -64..0 (flags 1) or 64 alone (flags 2) over
 * slime chunks it is the reference's filter. */
static int32_t floor_quart(int32_t y) { return y >= 0 ? y / 4 : -((3 - y) / 4); }

int ob_filter_levels(const ob_climate* c, int64_t world_seed, int32_t cx, int32_t cz, int flags,
                     int32_t y_min, int32_t y_max, int32_t y_step, int tree) {
    int want = flags & 3, mask = 0;
    for (int ci = 0; ci < 17 * 17; ++ci) {
        int dx = ci / 17 - 8, dz = ci % 17 - 8;
        if (dx * dx + dz * dz > 68) continue;
        int32_t x = (int32_t)((uint32_t)cx + (uint32_t)dx), z = (int32_t)((uint32_t)cz + (uint32_t)dz);
        if (!(flags & 4) && !slime(world_seed, x, z)) continue;
        for (int q = 0; q < 16; ++q) {
            if ((mask & want) == want) return mask;
            int32_t col[6];
            ob_column(c, (int32_t)((uint32_t)x * 4u + (uint32_t)(q / 4)),
                      (int32_t)((uint32_t)z * 4u + (uint32_t)(q % 4)), col);
            for (int64_t y = y_min; y <= y_max; y += y_step) {
                int b = biome_from_column(col, floor_quart((int32_t)y), tree);
                if ((want & 1) && b == 183) mask |= 1;
                if ((want & 2) && b == 14) mask |= 2;
            }
        }
    }
    return mask;
}
