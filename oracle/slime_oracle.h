/*
 * CPU oracle for the slime-chunk cluster search.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (slime_amd/, include/)
 * links, loads or calls this code. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 *
 * Each function is a scalar, from-scratch restatement of the reference
 * algorithm and cites the reference file:line it follows. Parity pinning is
 * described in oracle/README.md.
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

#ifdef __cplusplus
}
#endif

#endif
