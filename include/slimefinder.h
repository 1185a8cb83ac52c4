/*
 * slimefinder: MI355X (gfx950) slime-chunk cluster search, C ABI.
 *
 * Drop-in for the reference's native search libraries: SlimeCore.cpp (CPU)
 * and SlimeCoreGPU.cu (CUDA "v1"). Every function keeps the reference name,
 * argument meaning, result layout and error convention; the comment on each
 * cites the reference declaration it replaces. Plain C types only; the
 * library owns all device memory and the caller owns every host buffer.
 *
 * Searches are synchronous and run on the current HIP device. The control
 * functions are process-global and may be called from another thread while a
 * search runs, as in the reference.
 */
#ifndef SLIMEFINDER_H
#define SLIMEFINDER_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result row, 24 bytes: SlimeCore.cpp:182-189, SlimeCoreGPU.cu:198-205.
 * center_x/center_z are block coordinates cx*16+8 (wrapped to int32). With
 * precise_afk == 0, obs_count is 0 and afk_x/afk_z equal the center. */
typedef struct {
    int32_t size;
    int32_t center_x;
    int32_t center_z;
    int32_t obs_count;
    int32_t afk_x;
    int32_t afk_z;
} ExtChunkResult;

/* ---- search (returns the total number of centers with min_size <= count
 * <= max_size; fills the best min(total, max_results) rows ranked by size
 * desc, center_x asc, center_z asc; -1 on invalid arguments or device error) */

/* precise_afk != 0 refines the best rows after ranking (see the refinement
 * section below for the count): the CPU names follow SlimeCore.cpp:571-629
 * (the first 100 rows with the y scan enabled, else 500; the CPU rule; with
 * the y scan every row is packed, rows past the limit as (0, platform y)),
 * the GPU names SlimeCoreGPU.cu:3875-3912 (the first 100 rows; with the y
 * scan the GPU rule and packed, else the CPU rule unpacked). */

/* SlimeCore.cpp:643-651. `threads` is accepted and ignored (the GPU runs it). */
int64_t search_slime_clusters_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                                       int32_t max_size, int64_t rd_min, int32_t search_center_x,
                                       int32_t search_center_z, ExtChunkResult* results_buffer,
                                       int32_t max_results, int32_t threads, int32_t precise_afk);

/* SlimeCore.cpp:634-641: search around chunk (0, 0). */
int64_t search_slime_clusters(int64_t seed, int64_t rd_max, int32_t min_size, int32_t max_size,
                              int64_t rd_min, ExtChunkResult* results_buffer, int32_t max_results,
                              int32_t threads, int32_t precise_afk);

/* SlimeCoreGPU.cu:3927-3935 (GPU twin: no `threads`, rd_max <= (INT32_MAX-19)/2). */
int64_t search_slime_clusters_gpu_v1_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                                              int32_t max_size, int64_t rd_min,
                                              int32_t search_center_x, int32_t search_center_z,
                                              ExtChunkResult* results_buffer, int32_t max_results,
                                              int32_t precise_afk);

/* SlimeCoreGPU.cu:3918-3925. */
int64_t search_slime_clusters_gpu_v1(int64_t seed, int64_t rd_max, int32_t min_size,
                                     int32_t max_size, int64_t rd_min,
                                     ExtChunkResult* results_buffer, int32_t max_results,
                                     int32_t precise_afk);

/* SlimeCoreGPU.cu:3937-3955: legacy aliases. */
int64_t search_slime_clusters_gpu_v34(int64_t seed, int64_t rd_max, int32_t min_size,
                                      int32_t max_size, int64_t rd_min,
                                      ExtChunkResult* results_buffer, int32_t max_results,
                                      int32_t precise_afk);
int64_t search_slime_clusters_gpu_v34_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                                               int32_t max_size, int64_t rd_min,
                                               int32_t search_center_x, int32_t search_center_z,
                                               ExtChunkResult* results_buffer,
                                               int32_t max_results, int32_t precise_afk);

/* ---- control plane: SlimeCore.cpp:341-371, SlimeCoreGPU.cu:3085-3121 ---- */

void request_cancel(void);  /* also clears pause */
void request_pause(void);
void resume_search(void);
void reset_cancel(void);    /* also zeroes progress and the counters below */
int32_t get_progress(void); /* 1..79 while scanning, 90 when done */
int64_t get_processed_centers(void);
int64_t get_gpu_scan_work_ns(void);
void set_y_scan_config(int32_t enabled, int32_t platform_y, int32_t y_min, int32_t y_max,
                       int32_t y_step);

/* ---- precise AFK refinement: SlimeCore.cpp:654-773, SlimeCoreGPU.cu:3958-3990 ----
 * Both work on the caller's rows in place, on the GPU, with the heights of the
 * current set_y_scan_config. A row's AFK points are its center plus dx, dz in
 * -16..16 step 4 (scan_xz != 0; else the center only); obs_count is the number
 * of spawnable block columns of slime chunks at 24 < distance <= 128 from the
 * best point, packed as obs | (y + 1024) << 20 where noted.
 *
 * score_candidates_precise (SlimeCore.cpp:654-726; `threads` is accepted and
 * ignored): the CPU library's rule - the best starts at (0, center, platform
 * y) and only a strictly larger count replaces it; packed when the y scan is
 * enabled. */
void score_candidates_precise(int64_t seed, ExtChunkResult* results_buffer, int32_t count,
                              int32_t threads, int32_t scan_xz);
/* refine_candidates_y (SlimeCoreGPU.cu:3958-3990; the CPU library's version,
 * SlimeCore.cpp:728-773, adds a `threads` argument, accepted and ignored).
 * By default the GPU library's rule - equal counts go to the first (point,
 * height) in order, so a row whose counts are all zero gets point (-16, -16)
 * and the first height; with SLIME_REFINE_RULE=cpu in the environment (read
 * at every call) the CPU library's - the best starts at (0, center, platform
 * y) and only a strictly larger count replaces it. Always packed. */
void refine_candidates_y(int64_t seed, ExtChunkResult* results_buffer, int32_t count,
                         int32_t threads);

/* Extension with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): refine_candidates_y without allocating
 * anything. The rows are staged in a static device array of at most 65,536
 * rows (1.5 MiB, held by each device the library runs on until the process
 * ends; its address, a stream and two events are set up at the first call on
 * each device). Rows and result as refine_candidates_y. Returns 0, -1 on a
 * null buffer, negative count or device error (rows unchanged), -2 if count
 * exceeds 65,536. Calls in one process are serialized by a lock; the array
 * keeps the last call's rows. */
int32_t refine_candidates_y_noalloc(int64_t seed, ExtChunkResult* results_buffer, int32_t count);

/* Extension with no counterpart in the reference: refine_candidates_y over
 * the rows of many worlds in one launch. World w has seed world_seeds[w] and
 * rows [first[w], first[w+1]) of results_buffer (first[0] == 0); each row
 * gets exactly what refine_candidates_y gives it alone. 0, or -1 on invalid
 * arguments or a device error (rows unchanged). */
int32_t refine_candidates_y_multi(int32_t nworlds, const int64_t* world_seeds, const int32_t* first,
                                  ExtChunkResult* results_buffer);

/* SlimeCoreGPU.cu:3142-3161 (host-side single-chunk test). */
bool is_slime_chunk_c(int64_t seed, int64_t cx, int64_t cz);

/* SlimeCoreGPU.cu:3035-3083, 3118-3121, under HIP semantics. */
int32_t get_cuda_device_count(void);
int32_t get_cuda_device_name(char* buffer, int32_t buffer_size);
void cleanup_gpu_resources(void);
/* Last tile shape in the reference's numbering (1 = 128x8, 2 = 256x4,
 * 3 = 256x8, 4 = 512x4; 0 before any scan): 2 for the plain kernel, 3 for
 * the banded kernel. */
int32_t get_gpu_v1_shape(void);
/* The kernel that decided most centers of the last search, -1 before any:
 * 16 plain per-chunk, 17 banded one x-phase, 18 banded two x-phases (the
 * reference's ids 0-11 name its CUDA kernels). */
int32_t get_gpu_v1_algorithm(void);
/* The reference selects among three LCG formulations (0 native, 1 32-bit
 * limbs, 2 truncated); every kernel here computes the same exact test, so
 * these report the override set (0 when none) and change nothing else. */
int32_t get_gpu_v1_rng(void);
void set_gpu_v1_rng_override(int32_t value);
int32_t get_gpu_v34_shape(void);
int32_t get_gpu_v34_rng(void);
void set_gpu_v34_rng_override(int32_t value);

/* ---- extension with no counterpart in the reference ---------------------
 * Rectangle scan with the score histogram, for callers that partition a
 * region themselves (and for size-independent parity checks). Centers
 * [x0, x0+w) x [z0, z0+h); hist222 (optional) receives the count of in-range
 * hits per exact score 0..221. Same return convention as the searches. */
int64_t slime_scan_rect(int64_t seed, int32_t x0, int32_t z0, int64_t w, int64_t h,
                        int32_t min_size, int32_t max_size, ExtChunkResult* results_buffer,
                        int32_t max_results, int64_t* hist222);

/* Extension with no counterpart in the reference: the same scan over the
 * rows of [z0, z0+h) or, when band_rows > 0, only over the strided bands
 * z0 + b*2^17 + s (s < band_rows), with a forced scan mode: 0 automatic,
 * 1 plain per-chunk scan, 2 banded scan (one x-phase up to 2^17 columns,
 * two above), 3 banded with the square bound before refinement, 4 banded
 * with the union bound (the square minus cells outside the union of disks;
 * modes 0 and 2 take it), 5 as 4 with three x-phases (x, x + 2^17,
 * x + 2^18) over the columns past the last whole 2^18-column block of a
 * region wider than 2^18 (at least two bands). Lets tests and callers pick
 * the kernel explicitly. */
int64_t slime_scan_bands(int64_t seed, int32_t x0, int32_t z0, int64_t w, int64_t h,
                         int32_t band_rows, int32_t mode, int32_t min_size, int32_t max_size,
                         ExtChunkResult* results_buffer, int32_t max_results, int64_t* hist222);

/* Extension with no counterpart in the reference: device time of the scan
 * kernel launches since the last reset, measured with HIP events on the
 * engine's stream, for one kernel kind (0 plain per-chunk scan, 1 banded
 * scan, 2 both, 3 the banded launches with three x-phases, a part of 1),
 * with the number of candidate centers those launches decided. reset != 0
 * zeroes every counter after the read. */
void slime_kernel_stats(int32_t kind, int64_t* launches, int64_t* kernel_ns, int64_t* centers,
                        int32_t reset);

/* Extension with no counterpart in the reference: compute units, peak shader
 * clock (kHz) and HBM size of the current device; 0 on success, -1 if no
 * device is usable. */
int32_t slime_device_info(int32_t* compute_units, int32_t* clock_khz, int64_t* hbm_bytes);

/* Extension with no counterpart in the reference (measurement only): the
 * device's issue rate of one kind of 32-bit integer VALU instruction, the
 * measured ceiling the scan's roofline is read against. Runs iters x 64
 * instructions of the kind in 8 independent chains per lane, 8 workgroups of
 * 256 threads per compute unit; kind 0 v_add_u32, 1 v_xor_b32 (VOP2, 4
 * bytes), 2 v_mul_lo_u32, 3 v_bcnt_u32_b32, 4 v_mad_u32_u24, 5 v_alignbit_b32
 * (VOP3, 8 bytes), 6 v_xor_b32 with a 32-bit literal (8 bytes), 7 v_add_u32 in
 * its VOP3 encoding (8 bytes), 8 v_mul_u32_u24 (VOP2), 9 v_xor_b32 with an
 * SGPR operand (VOP2).
 * Writes the lane instructions executed and the kernel's device time (HIP
 * events); 0 on success, -1 on bad arguments (iters in 1..2^20) or a device
 * error. */
int32_t slime_valu_rate(int32_t kind, int32_t iters, int64_t* lane_instructions,
                        int64_t* kernel_ns);

/* Extension with no counterpart in the reference (whose binary tensor-core
 * popcounts are cuda_x2_z15/tensor_cold.py:4-39); measurement only. A/B of
 * the banded scan's counting stage over a hashed bitplane ring in LDS laid
 * out as the scan's: the rolling 20x20 window of every 4x4 group less its 32
 * corner cells, per band-phase, for `steps` strips of 4 rows. method 0 counts
 * on the VALU by plane words as the scan kernels do; method 1 forms the
 * 5-column window sums with int8 matrix instructions (v_mfma_i32_16x16x64_i8).
 * bands 2, 4 or 8 (two x-phases each), blocks 1..65536 workgroups, steps
 * 5..2^20, min_size 1..400. Writes the wrapping sum of a weight of (strip,
 * band-phase, group) over the windows reaching min_size, their number, and
 * the timed launch's device time (HIP events); 0 on success, -1 on bad
 * arguments or a device error. */
int32_t slime_window_ab(int32_t method, int32_t bands, int32_t blocks, int32_t steps,
                        int32_t min_size, uint32_t seed, uint32_t* checksum, uint64_t* passes,
                        int64_t* kernel_ns);

/* Extension with no counterpart in the reference: device time (ns, HIP
 * events) of the last precise-AFK refinement kernel and its number of
 * (row, AFK point, height within 128 blocks of the platform) evaluations. */
void slime_refine_stats(int64_t* kernel_ns, int64_t* evaluations);

/* ---- batched searches (extension; the reference runs one search per call)
 * n independent searches, each with the meaning of
 * search_slime_clusters_gpu_v1_centered (SlimeCoreGPU.cu:3927-3935), scanned
 * together: one launch per kernel class over all the searches' tiles (one
 * workgroup per tile), or with flags bit 1 one persistent launch per kernel
 * class fed from a device work queue (per-job claim cursors; the form served
 * sessions use). Search i writes up to max_results ranked rows at
 * results_buffer + i * max_results and its return value (hit count, or -1 for
 * invalid arguments) to totals[i]. flags bit 0: replay the launches from a
 * captured hipGraph. Returns 0, or -1 on bad arrays or a device error.
 * Results are identical to n separate calls. */
int32_t slime_search_batch(int32_t n, const int64_t* seeds, const int64_t* rd_max,
                           const int32_t* min_size, const int32_t* max_size, const int64_t* rd_min,
                           const int32_t* center_x, const int32_t* center_z,
                           ExtChunkResult* results_buffer, int32_t max_results, int64_t* totals,
                           int32_t flags);

/* Extension: figures of the last batch: candidate centers, device time of the
 * sweep launches, host time of planning + fix-up + upload, of the sweep, and
 * of the whole call (ns), kernel launches, jobs, searches rerun alone after a
 * hit-region overflow, hits stored in the searches' device regions and how
 * many of them came back to the host to be ranked (a search's best rows are
 * selected on the device first: only the hits at or above the score its
 * top max_results reach). Any pointer may be NULL. */
void slime_batch_stats(int64_t* centers, int64_t* kernel_ns, int64_t* prepare_ns,
                       int64_t* sweep_ns, int64_t* wall_ns, int32_t* launches, int32_t* jobs,
                       int32_t* fallbacks, int64_t* hits_stored, int64_t* hits_copied);

/* Extension with no counterpart in the reference: the centers the last
 * slime_search_batch call decided per kernel class (the order of
 * slime_sweep_costs). Writes at most cap entries; returns how many, or -1 for
 * a null buffer. */
int32_t slime_batch_class_centers(int64_t* centers, int32_t cap);

/* Extension: the autotuner's measured cost table, one entry per kernel class
 * (bands nb, x-phases ph): device ns per tile row with the whole GPU on that
 * class, and the fixed per-tile cost in rows. Measures on first use. Returns
 * the number of classes written (at most cap), or -1 on a device error. */
int32_t slime_sweep_costs(double* ns_row, double* extra_rows, int32_t* nb, int32_t* ph,
                          int32_t cap);

/* ---- served searches (extension; the reference serves one search at a time
 * behind process-global flags, SlimeCore.cpp:16-27) -------------------------
 * A session keeps every admitted search's state on the device (jobs, per-job
 * done counts and claim cursors, histogram, hit region, status word) and scans
 * all of them together in rounds (continuous batching): per kernel class one
 * launch over the tiles of every active request, least work left first.
 * Requests are admitted in arrival order while their device footprint fits
 * hbm_budget bytes and fewer than max_active are admitted; round_ms bounds the
 * device time of one round (the latency of admission, pause and cancel). Each
 * result equals search_slime_clusters_gpu_v1_centered with the same
 * arguments; a cancelled request keeps the exact rows of the tiles already
 * scanned. Request states: 0 queued, 1 running, 2 paused, 3 done, 4 cancelled
 * (partial result), 5 failed. A request over this build's size limit (2^48
 * centers: radius 8,388,607), over the budget alone, or with more jobs or
 * hits than the session holds is refused at submit (-1).
 * slime_session_create runs direct rounds (one workgroup per tile, the
 * per-search kernels' register budget; pause and cancel act between rounds),
 * pipelined: the next round is queued shortly before the current one ends, so
 * a pause or cancel takes effect after at most the round in flight and short
 * rounds (10 ms) cost little throughput;
 * slime_session_create_ex with flags bit 0 runs the persistent work-queue
 * kernels instead, which read each request's status word before every tile
 * claim. NULL on bad arguments or no device. */
void* slime_session_create(int64_t hbm_budget, int32_t max_active, int32_t round_ms);
void* slime_session_create_ex(int64_t hbm_budget, int32_t max_active, int32_t round_ms,
                              int32_t flags);
void slime_session_destroy(void* session);
/* Returns a request id (> 0), or -1 for invalid arguments. */
int64_t slime_session_submit(void* session, int64_t seed, int64_t rd_max, int32_t min_size,
                             int32_t max_size, int64_t rd_min, int32_t search_center_x,
                             int32_t search_center_z, int32_t max_results);
/* 0 on success, -1 if the request is unknown or already finished. */
int32_t slime_session_pause(void* session, int64_t id);
int32_t slime_session_resume(void* session, int64_t id);
int32_t slime_session_cancel(void* session, int64_t id);
/* Returns the state (or -1 for an unknown id); progress in percent, total hits
 * (-1 until done) and submit-to-finish latency in ns (-1 until done). */
int32_t slime_session_poll(void* session, int64_t id, int32_t* progress_pct, int64_t* total,
                           int64_t* latency_ns);
/* Blocks until the request finishes or timeout_ms passes (< 0: no limit);
 * returns its state. */
int32_t slime_session_wait(void* session, int64_t id, int32_t timeout_ms);
/* Copies up to cap ranked rows of a finished request; returns its total, or
 * -1 if it is not finished. */
int64_t slime_session_result(void* session, int64_t id, ExtChunkResult* results_buffer,
                             int32_t cap);
/* Forgets a finished request; 0 on success. */
int32_t slime_session_release(void* session, int64_t id);
/* Session figures: centers of completed requests, device time, rounds,
 * admitted and queued requests, device bytes held and the budget. */
void slime_session_stats(void* session, int64_t* centers_done, int64_t* kernel_ns,
                         int32_t* rounds, int32_t* active, int32_t* queued, int64_t* hbm_in_use,
                         int64_t* hbm_budget);

#ifdef __cplusplus
}
#endif

#endif /* SLIMEFINDER_H */
