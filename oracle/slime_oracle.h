/*
 * CPU oracle for the slime-chunk cluster search.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (slime_amd/, include/)
 * links, loads or calls this code. Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline legs use it, as the checker and the timed CPU
 * baseline.
 *
 * Each function is a scalar, from-scratch restatement of the reference
 * algorithm and cites the reference file:line it follows. The tests under
 * tests/ pin each one against the compiled reference, known answers or a
 * brute force.
 */
#ifndef SLIME_ORACLE_H
#define SLIME_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t size;
    int32_t center_x;
    int32_t center_z;
    int32_t obs_count;
    int32_t afk_x;
    int32_t afk_z;
} OracleResult;

/* Java-Random slime predicate (SlimeCore.cpp:208-232). */
int32_t oracle_is_slime(int64_t seed, int64_t cx, int64_t cz);

/* First-draw-only predicate: (next(31) % 10 == 0) without the rejection redraw.
 * Used by tests of the first-draw scan plus its rejection correction. */
int32_t oracle_first_draw(int64_t seed, int64_t cx, int64_t cz);

/* Java nextInt(10)==0 evaluated from a raw 48-bit LCG state taken as the
 * state after the scramble (SlimeCore.cpp:68-80 applies it after one step). */
int32_t oracle_slime_from_state0(uint64_t state0);

/* new java.util.Random(seed).nextInt(): known-answer hook for the LCG. */
int32_t oracle_java_next_int(int64_t seed);

/* Coordinate terms (SlimeCore.cpp:442-449): x-term and z-term, mod 2^64. */
uint64_t oracle_x_term(int64_t cx);
uint64_t oracle_z_term(int64_t cz);

/* Exact 221-cell disk count around one chunk center. */
int32_t oracle_disk_count(int64_t seed, int64_t cx, int64_t cz);

/* Number of scrambled seeds s (mod 2^48) whose first draw is rejected and
 * whose redraw chain ends on a multiple of 10
 * (cuda_x2_z15/generate_slime_rare_tables.py:15-33 fixes it at 104,848). */
int64_t oracle_count_rejected_hits(void);

/* Drop-in search (SlimeCore.cpp:373-651), precise_afk == 0 only.
 * Returns the total hit count, or -1 on invalid arguments. */
int64_t oracle_search_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                               int32_t max_size, int64_t rd_min, int32_t center_x,
                               int32_t center_z, OracleResult* results,
                               int32_t max_results);

/* Same search with a score histogram (hist[0..221]) of every hit, used by the
 * tests of the device-side histogram and top-K. */
int64_t oracle_search_hist(int64_t seed, int64_t rd_max, int32_t min_size,
                           int32_t max_size, int64_t rd_min, int32_t center_x,
                           int32_t center_z, OracleResult* results,
                           int32_t max_results, int64_t* hist222);

/* Rectangle search over centers [x0, x0+w) x [z0, z0+h): used to time the
 * CPU baseline on bounded samples and by window checks. */
int64_t oracle_search_rect(int64_t seed, int64_t x0, int64_t z0, int64_t w,
                           int64_t h, int32_t min_size, int32_t max_size,
                           OracleResult* results, int32_t max_results,
                           int64_t* hist222);

/* Spawnable block columns around an AFK point (SlimeCore.cpp:234-289,
 * count_spawnable_cached): the blocks (x, z) of slime chunks with
 * 576 < (x-ox)^2 + (z-oz)^2 + dy^2 <= 16384, dy = platform_y - afk_y. */
int32_t oracle_spawnable(int64_t seed, int64_t ox, int64_t oz, int32_t dy);

/* Precise AFK refinement of result rows in place (SlimeCore.cpp:578-629,
 * 654-726; SlimeCoreGPU.cu:2749-2892). Offsets: 81 points (dx, dz in
 * -16..16 step 4, dx outer) when scan_xz, else the center only. Y levels
 * y_start + i*y_step for i < y_count, evaluated against platform_y. rule 0:
 * best starts at (0, center, platform_y) and only a strictly larger count
 * replaces it (the CPU loops and refine_afk_kernel); rule 1: best starts at
 * -1 and ties go to the lowest (offset, level) order (refine_afk_block_kernel,
 * refine_afk_xz_kernel). pack != 0 stores obs | (y + 1024) << 20. */
void oracle_refine_rows(int64_t seed, OracleResult* rows, int32_t count, int32_t scan_xz,
                        int32_t platform_y, int32_t y_start, int32_t y_step, int64_t y_count,
                        int32_t rule, int32_t pack);

#ifdef __cplusplus
}
#endif

#endif
