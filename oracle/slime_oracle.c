/*
 * CPU oracle for the slime-chunk cluster search (scalar C, one thread).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path. The product never
 * links or calls it. It restates the reference algorithm from the Java
 * Random specification and the reference's search semantics:
 *   - slime predicate ............ SlimeCore.cpp:208-232 (checkSlime),
 *                                  SlimeCore.cpp:54-80 (nextInt(10) rejection)
 *   - disk mask (dx^2+dz^2 <= 68)  SlimeCore.cpp:399-404
 *   - per-center exact count ..... SlimeCore.cpp:526-533
 *   - argument validation ........ SlimeCore.cpp:378-384
 *   - rd_min annulus ............. SlimeCore.cpp:504-509
 *   - ranking (size desc, x asc, z asc), block coords cx*16+8
 *                                  SlimeCore.cpp:191-206, 535-569
 *   - result rows with precise_afk == 0 (obs 0, afk = center)
 *                                  SlimeCore.cpp:580-628
 * The algorithmic shape is deliberately different from the reference (row
 * prefix sums over strips, one binary heap), so a shared mistake is unlikely.
 */
#include "slime_oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define MASK48 0xFFFFFFFFFFFFULL
#define JAVA_MUL 0x5DEECE66DULL
#define JAVA_ADD 0xBULL
#define SLIME_XOR 0x3AD8025FULL

/* Half-width of the disk row at |dz| (largest d with dz^2 + d^2 <= 68). */
static int32_t disk_half_width(int32_t dz) {
    int32_t d = 0;
    while (dz * dz + (d + 1) * (d + 1) <= 68) d++;
    return d;
}

uint64_t oracle_x_term(int64_t cx) {
    /* Java int arithmetic: (int)(x*x*4987142) + (int)(x*5947611), widened. */
    uint32_t u = (uint32_t)(int32_t)cx;
    int32_t a = (int32_t)(u * u * 4987142u);
    int32_t b = (int32_t)(u * 5947611u);
    return (uint64_t)((int64_t)a + (int64_t)b);
}

uint64_t oracle_z_term(int64_t cz) {
    /* Java: (long)(z*z) * 4392871L + (int)(z*389711). */
    uint32_t u = (uint32_t)(int32_t)cz;
    int64_t sq = (int64_t)(int32_t)(u * u);
    int32_t lin = (int32_t)(u * 389711u);
    return (uint64_t)(sq * 4392871LL) + (uint64_t)(int64_t)lin;
}

/* java.util.Random.next(31) on a 48-bit state; advances the state. */
static uint32_t java_next31(uint64_t* state) {
    *state = (*state * JAVA_MUL + JAVA_ADD) & MASK48;
    return (uint32_t)(*state >> 17);
}

int32_t oracle_java_next_int(int64_t seed) {
    uint64_t st = ((uint64_t)seed ^ JAVA_MUL) & MASK48;
    st = (st * JAVA_MUL + JAVA_ADD) & MASK48;
    return (int32_t)(uint32_t)(st >> 16);
}

/* nextInt(10) == 0 with the rejection loop, from the scrambled state. */
int32_t oracle_slime_from_state0(uint64_t state0) {
    uint64_t st = state0 & MASK48;
    for (;;) {
        int32_t bits = (int32_t)java_next31(&st);
        int32_t val = bits % 10;
        /* Java rejects when bits - val + 9 overflows int32. */
        if ((int64_t)bits - val + 9 <= INT32_MAX) return val == 0;
    }
}

static uint64_t scrambled_state(int64_t seed, int64_t cx, int64_t cz) {
    uint64_t s = (uint64_t)seed + oracle_x_term(cx) + oracle_z_term(cz);
    return ((s ^ SLIME_XOR) ^ JAVA_MUL) & MASK48;
}

int32_t oracle_is_slime(int64_t seed, int64_t cx, int64_t cz) {
    return oracle_slime_from_state0(scrambled_state(seed, cx, cz));
}

int32_t oracle_first_draw(int64_t seed, int64_t cx, int64_t cz) {
    uint64_t st = scrambled_state(seed, cx, cz);
    uint32_t bits = java_next31(&st);
    return bits % 10u == 0;
}

int32_t oracle_disk_count(int64_t seed, int64_t cx, int64_t cz) {
    int32_t total = 0;
    for (int32_t dz = -8; dz <= 8; ++dz) {
        int32_t hw = disk_half_width(dz < 0 ? -dz : dz);
        for (int32_t dx = -hw; dx <= hw; ++dx)
            total += oracle_is_slime(seed, cx + dx, cz + dz);
    }
    return total;
}

int64_t oracle_count_rejected_hits(void) {
    /* Rejected first draws are exactly the states >= (2^31 - 8) << 17. */
    int64_t n = 0;
    const uint64_t lo = ((uint64_t)(INT32_MAX - 7)) << 17;
    for (uint64_t st1 = lo; st1 <= MASK48; ++st1) {
        uint64_t st = st1;
        int32_t bits = (int32_t)(st >> 17);
        int32_t val = bits % 10;
        while ((int64_t)bits - val + 9 > INT32_MAX) {
            bits = (int32_t)java_next31(&st);
            val = bits % 10;
        }
        n += (val == 0);
    }
    return n;
}

/* ---- top-K heap: root is the worst kept entry ------------------------- */

typedef struct {
    int32_t size;
    int64_t bx, bz; /* block coordinates cx*16+8, compared as int64 */
} Entry;

/* a ranks before b */
static int better(const Entry* a, const Entry* b) {
    if (a->size != b->size) return a->size > b->size;
    if (a->bx != b->bx) return a->bx < b->bx;
    return a->bz < b->bz;
}

typedef struct {
    Entry* v;
    int64_t n, cap;
} Heap;

static void heap_sift_down(Heap* h, int64_t i) {
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, w = i;
        if (l < h->n && better(&h->v[w], &h->v[l])) w = l;
        if (r < h->n && better(&h->v[w], &h->v[r])) w = r;
        if (w == i) return;
        Entry t = h->v[i]; h->v[i] = h->v[w]; h->v[w] = t;
        i = w;
    }
}

static void heap_offer(Heap* h, Entry e) {
    if (h->n < h->cap) {
        int64_t i = h->n++;
        h->v[i] = e;
        while (i > 0) {
            int64_t p = (i - 1) / 2;
            if (!better(&h->v[p], &h->v[i])) break;
            Entry t = h->v[i]; h->v[i] = h->v[p]; h->v[p] = t;
            i = p;
        }
    } else if (better(&e, &h->v[0])) {
        h->v[0] = e;
        heap_sift_down(h, 0);
    }
}

static int cmp_entry(const void* a, const void* b) {
    const Entry* x = (const Entry*)a;
    const Entry* y = (const Entry*)b;
    if (better(x, y)) return -1;
    if (better(y, x)) return 1;
    return 0;
}

/* Scan centers [x0,x0+w) x [z0,z0+h); annulus test against (acx,acz) when
 * rd_min > 0. Returns total hits; fills heap and hist. */
static int64_t scan_rect(int64_t seed, int64_t x0, int64_t z0, int64_t w, int64_t h,
                         int32_t min_size, int32_t max_size, int64_t rd_min,
                         int64_t acx, int64_t acz, Heap* heap, int64_t* hist) {
    const int64_t STRIP = 64;
    const int64_t cols = w + 16; /* chunk columns x0-8 .. x0+w+7 */
    int32_t hw[17];
    for (int32_t dz = -8; dz <= 8; ++dz) hw[dz + 8] = disk_half_width(dz < 0 ? -dz : dz);
    int32_t* pre = (int32_t*)malloc(sizeof(int32_t) * (size_t)(STRIP + 16) * (size_t)(cols + 1));
    uint64_t* xt = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)cols);
    if (!pre || !xt) { free(pre); free(xt); return -1; }
    for (int64_t i = 0; i < cols; ++i) xt[i] = (uint64_t)seed + oracle_x_term(x0 - 8 + i);
    const int64_t rmin2 = rd_min * rd_min;
    int64_t total = 0;
    for (int64_t sz = 0; sz < h; sz += STRIP) {
        int64_t sh = h - sz < STRIP ? h - sz : STRIP;
        int64_t rows = sh + 16;
        for (int64_t r = 0; r < rows; ++r) {
            int64_t cz = z0 + sz - 8 + r;
            uint64_t zt = oracle_z_term(cz);
            int32_t* p = pre + r * (cols + 1);
            p[0] = 0;
            for (int64_t i = 0; i < cols; ++i) {
                uint64_t st0 = ((xt[i] + zt) ^ SLIME_XOR ^ JAVA_MUL) & MASK48;
                p[i + 1] = p[i] + oracle_slime_from_state0(st0);
            }
        }
        for (int64_t j = 0; j < sh; ++j) {
            int64_t cz = z0 + sz + j;
            for (int64_t i = 0; i < w; ++i) {
                int64_t cx = x0 + i;
                if (rd_min > 0) {
                    int64_t dx = cx - acx, dz = cz - acz;
                    if (dx * dx + dz * dz < rmin2) continue;
                }
                int32_t cnt = 0;
                for (int32_t dz = -8; dz <= 8; ++dz) {
                    const int32_t* p = pre + (j + 8 + dz) * (cols + 1);
                    int64_t c = i + 8; /* column of cx */
                    cnt += p[c + hw[dz + 8] + 1] - p[c - hw[dz + 8]];
                }
                if (cnt < min_size || cnt > max_size) continue;
                total++;
                if (hist) hist[cnt]++;
                if (heap) {
                    Entry e = {cnt, cx * 16 + 8, cz * 16 + 8};
                    heap_offer(heap, e);
                }
            }
        }
    }
    free(pre);
    free(xt);
    return total;
}

static void write_results(Heap* heap, OracleResult* out) {
    qsort(heap->v, (size_t)heap->n, sizeof(Entry), cmp_entry);
    for (int64_t i = 0; i < heap->n; ++i) {
        out[i].size = heap->v[i].size;
        out[i].center_x = (int32_t)heap->v[i].bx;
        out[i].center_z = (int32_t)heap->v[i].bz;
        out[i].obs_count = 0;
        out[i].afk_x = (int32_t)heap->v[i].bx;
        out[i].afk_z = (int32_t)heap->v[i].bz;
    }
}

int64_t oracle_search_hist(int64_t seed, int64_t rd_max, int32_t min_size,
                           int32_t max_size, int64_t rd_min, int32_t center_x,
                           int32_t center_z, OracleResult* results,
                           int32_t max_results, int64_t* hist222) {
    if (!results || max_results <= 0 || rd_max < 0 ||
        rd_max > (INT32_MAX - 17LL) / 2LL || rd_min < 0 || rd_min > rd_max ||
        min_size < 0 || min_size > 221 || max_size < min_size)
        return -1;
    if ((int64_t)center_x - rd_max - 8 < INT32_MIN || (int64_t)center_x + rd_max + 8 > INT32_MAX ||
        (int64_t)center_z - rd_max - 8 < INT32_MIN || (int64_t)center_z + rd_max + 8 > INT32_MAX)
        return -1;
    Heap heap;
    heap.n = 0;
    heap.cap = max_results;
    heap.v = (Entry*)malloc(sizeof(Entry) * (size_t)max_results);
    if (!heap.v) return -1;
    if (hist222) memset(hist222, 0, sizeof(int64_t) * 222);
    int64_t total = scan_rect(seed, (int64_t)center_x - rd_max, (int64_t)center_z - rd_max,
                              2 * rd_max + 1, 2 * rd_max + 1, min_size, max_size, rd_min,
                              center_x, center_z, &heap, hist222);
    if (total >= 0) write_results(&heap, results);
    free(heap.v);
    return total;
}

int64_t oracle_search_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                               int32_t max_size, int64_t rd_min, int32_t center_x,
                               int32_t center_z, OracleResult* results,
                               int32_t max_results) {
    return oracle_search_hist(seed, rd_max, min_size, max_size, rd_min, center_x, center_z,
                              results, max_results, NULL);
}

int64_t oracle_search_rect(int64_t seed, int64_t x0, int64_t z0, int64_t w, int64_t h,
                           int32_t min_size, int32_t max_size, OracleResult* results,
                           int32_t max_results, int64_t* hist222) {
    if (w <= 0 || h <= 0 || min_size < 0 || max_size < min_size) return -1;
    Heap heap;
    heap.n = 0;
    heap.cap = max_results > 0 ? max_results : 1;
    heap.v = (Entry*)malloc(sizeof(Entry) * (size_t)heap.cap);
    if (!heap.v) return -1;
    if (hist222) memset(hist222, 0, sizeof(int64_t) * 222);
    int64_t total = scan_rect(seed, x0, z0, w, h, min_size, max_size, 0, 0, 0,
                              results ? &heap : NULL, hist222);
    if (total >= 0 && results) write_results(&heap, results);
    free(heap.v);
    return total;
}

/* ---- precise AFK refinement ------------------------------------------------
 * Restates SlimeCore.cpp:148-180 (radius tables), :234-289 (spawnable count),
 * :578-629 and :654-726 (offset / level loops) and the GPU selection rule of
 * SlimeCoreGPU.cu:2799-2861. Shape chosen differently from the reference: a
 * 25x25 chunk bitmap per row, then per x column the ring's z intervals walked
 * chunk by chunk. */

static int32_t isqrt_exact(int64_t v) {
    int64_t r = 0;
    while ((r + 1) * (r + 1) <= v) r++;
    return (int32_t)r;
}

static int64_t floor16(int64_t v) { return v >= 0 ? v / 16 : -((15 - v) / 16); }

typedef struct {
    int64_t cx0, cz0;     /* chunk of bitmap cell (0, 0) */
    uint8_t bit[25][25];  /* [x][z] */
} ChunkWindow;

static void window_fill(ChunkWindow* w, int64_t seed, int64_t bx, int64_t bz) {
    w->cx0 = floor16(bx) - 12;
    w->cz0 = floor16(bz) - 12;
    for (int i = 0; i < 25; ++i)
        for (int j = 0; j < 25; ++j)
            w->bit[i][j] = (uint8_t)oracle_is_slime(seed, w->cx0 + i, w->cz0 + j);
}

/* Blocks z in [a, b] of chunk column cx that lie in slime chunks. */
static int32_t column_run(const ChunkWindow* w, int64_t cx, int64_t a, int64_t b) {
    int32_t n = 0;
    for (int64_t cz = floor16(a); cz <= floor16(b); ++cz) {
        if (!w->bit[cx - w->cx0][cz - w->cz0]) continue;
        int64_t lo = a > cz * 16 ? a : cz * 16;
        int64_t hi = b < cz * 16 + 15 ? b : cz * 16 + 15;
        if (hi >= lo) n += (int32_t)(hi - lo + 1);
    }
    return n;
}

static int32_t window_spawnable(const ChunkWindow* w, int64_t ox, int64_t oz, int64_t dy) {
    if (dy * dy > 16384) return 0;
    int32_t total = 0;
    for (int64_t dx = -128; dx <= 128; ++dx) {
        int64_t outer2 = 16384 - dy * dy - dx * dx;
        if (outer2 < 0) continue;
        int64_t ro = isqrt_exact(outer2);
        int64_t inner2 = 576 - dy * dy - dx * dx;
        int64_t cx = floor16(ox + dx);
        if (inner2 < 0) {
            total += column_run(w, cx, oz - ro, oz + ro);
        } else {
            int64_t ri = isqrt_exact(inner2);
            total += column_run(w, cx, oz - ro, oz - ri - 1);
            total += column_run(w, cx, oz + ri + 1, oz + ro);
        }
    }
    return total;
}

int32_t oracle_spawnable(int64_t seed, int64_t ox, int64_t oz, int32_t dy) {
    ChunkWindow w;
    window_fill(&w, seed, ox, oz);
    return window_spawnable(&w, ox, oz, dy);
}

void oracle_refine_rows(int64_t seed, OracleResult* rows, int32_t count, int32_t scan_xz,
                        int32_t platform_y, int32_t y_start, int32_t y_step, int64_t y_count,
                        int32_t rule, int32_t pack) {
    ChunkWindow w;
    const int32_t span = scan_xz ? 16 : 0, step = scan_xz ? 4 : 1;
    for (int32_t r = 0; r < count; ++r) {
        const int64_t bx = rows[r].center_x, bz = rows[r].center_z;
        window_fill(&w, seed, bx, bz);
        int32_t best_obs = rule ? -1 : 0;
        int64_t best_x = bx, best_z = bz, best_y = platform_y;
        for (int32_t dx = -span; dx <= span; dx += step) {
            for (int32_t dz = -span; dz <= span; dz += step) {
                for (int64_t i = 0; i < y_count; ++i) {
                    const int64_t y = (int64_t)y_start + i * y_step;
                    const int64_t dy = (int64_t)platform_y - y;
                    /* Levels farther than 128 blocks count 0; skipping them
                     * keeps huge level lists cheap. */
                    const int32_t obs =
                        (dy > 128 || dy < -128) ? 0 : window_spawnable(&w, bx + dx, bz + dz, dy);
                    if (obs > best_obs) {
                        best_obs = obs;
                        best_x = bx + dx;
                        best_z = bz + dz;
                        best_y = y;
                    }
                }
            }
        }
        rows[r].obs_count =
            pack ? (int32_t)((uint32_t)best_obs | ((uint32_t)(best_y + 1024) << 20)) : best_obs;
        rows[r].afk_x = (int32_t)best_x;
        rows[r].afk_z = (int32_t)best_z;
    }
}
