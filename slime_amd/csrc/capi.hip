// C ABI of libslimefinder.so: the reference's search and control exports.
//
// Argument validation, return conventions and control semantics follow
// SlimeCore.cpp:340-651 (CPU names) and SlimeCoreGPU.cu:3034-3955 (GPU
// names); the two families differ in the rd_max limit ((INT32_MAX-17)/2 vs
// (INT32_MAX-19)/2), the CPU `threads` argument, and the precise-AFK
// selection rule and row limits (SlimeCore.cpp:571-629 vs
// SlimeCoreGPU.cu:3875-3912).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/slimefinder.h"
#include "engine.h"
#include "refine.h"
#include "slime_math.h"
#include "sweep_engine.h"

using slime::control;

static_assert(sizeof(ExtChunkResult) == sizeof(slime::ResultRow), "row layout");

namespace {

int32_t g_rng_override = -1;

struct YScan {
    int32_t enabled = 0, platform_y = -64, y_min = -64, y_max = 64, y_step = 4;
} g_yscan;

// SlimeCore.cpp:378-384 / SlimeCoreGPU.cu:3169-3175.
bool valid_args(int64_t rd_max, int64_t rd_max_limit, int32_t min_size, int32_t max_size,
                int64_t rd_min, int32_t cx, int32_t cz, const ExtChunkResult* out,
                int32_t max_results) {
    if (!out || max_results <= 0 || rd_max < 0 || rd_max > rd_max_limit || rd_min < 0 ||
        rd_min > rd_max || min_size < 0 || min_size > 221 || max_size < min_size)
        return false;
    if ((int64_t)cx - rd_max - 8 < INT32_MIN || (int64_t)cx + rd_max + 8 > INT32_MAX ||
        (int64_t)cz - rd_max - 8 < INT32_MIN || (int64_t)cz + rd_max + 8 > INT32_MAX)
        return false;
    // this build's size limit (radius 8,388,607; the reference's is larger)
    return slime::within_search_limit(2 * rd_max + 1, 2 * rd_max + 1);
}

// The square of centers within rd_max (Chebyshev) of (cx, cz), with the
// annulus filter rd_min (SlimeCore.cpp:430-449, 504-512).
slime::SearchSpec square_spec(int64_t seed, int64_t rd_max, int32_t min_size, int32_t max_size,
                              int64_t rd_min, int32_t cx, int32_t cz) {
    slime::SearchSpec spec;
    spec.seed = seed;
    spec.x0 = (int32_t)((int64_t)cx - rd_max);
    spec.z0 = (int32_t)((int64_t)cz - rd_max);
    spec.w = 2 * rd_max + 1;
    spec.h = 2 * rd_max + 1;
    spec.min_size = min_size;
    spec.max_size = max_size;
    spec.rd_min = rd_min;
    spec.acx = cx;
    spec.acz = cz;
    return spec;
}

int64_t search_common(int64_t seed, int64_t rd_max, int32_t min_size, int32_t max_size,
                      int64_t rd_min, int32_t cx, int32_t cz, ExtChunkResult* out,
                      int32_t max_results, int64_t rd_max_limit) {
    if (!valid_args(rd_max, rd_max_limit, min_size, max_size, rd_min, cx, cz, out, max_results))
        return -1;
    control().progress.store(1, std::memory_order_relaxed);
    const slime::SearchSpec spec = square_spec(seed, rd_max, min_size, max_size, rd_min, cx, cz);
    return slime::run_search(spec, reinterpret_cast<slime::ResultRow*>(out), max_results);
}

slime::RefineSpec refine_spec(int64_t seed, int32_t scan_xz, int32_t rule, int32_t pack) {
    slime::RefineSpec r;
    r.seed = seed;
    r.scan_xz = scan_xz;
    r.levels = slime::y_levels(g_yscan.enabled, g_yscan.platform_y, g_yscan.y_min, g_yscan.y_max,
                               g_yscan.y_step);
    r.rule = rule;
    r.pack = pack;
    return r;
}

bool cancelled() { return control().cancel.load(std::memory_order_relaxed) != 0; }

// precise_afk != 0 of the CPU names (SlimeCore.cpp:571-629): the first 100
// rows (y scan) or 500 rows are refined with the keep-center rule; with the y
// scan every row is packed, the unrefined ones as (0, platform_y). A device
// error leaves the rows as ranked.
void precise_cpu(int64_t seed, ExtChunkResult* out, int64_t total, int32_t max_results) {
    const int32_t n = (int32_t)std::min<int64_t>(total, max_results);
    if (n <= 0 || cancelled()) return;
    const int32_t scan_y = g_yscan.enabled;
    const int32_t limit = std::min(n, scan_y ? 100 : 500);
    auto* rows = reinterpret_cast<slime::ResultRow*>(out);
    if (slime::refine_rows(refine_spec(seed, 1, slime::kRuleKeepCenter, scan_y), rows, limit) != 0)
        return;
    if (scan_y)
        for (int32_t i = limit; i < n; ++i)
            out[i].obs_count = (int32_t)((uint32_t)(g_yscan.platform_y + 1024) << 20);
}

// precise_afk != 0 of the GPU names (SlimeCoreGPU.cu:3875-3912): the first
// 100 rows; with the y scan the block kernel's first-in-order rule, packed;
// else the thread kernel's keep-center rule, unpacked.
void precise_gpu(int64_t seed, ExtChunkResult* out, int64_t total, int32_t max_results) {
    const int32_t n = (int32_t)std::min<int64_t>(total, max_results);
    if (n <= 0 || cancelled()) return;
    const int32_t scan_y = g_yscan.enabled;
    (void)slime::refine_rows(
        refine_spec(seed, 1, scan_y ? slime::kRuleFirstOrder : slime::kRuleKeepCenter, scan_y),
        reinterpret_cast<slime::ResultRow*>(out), std::min(n, 100));
}

// refine_candidates_y is exported by both reference libraries, with different
// arity and a different rule for rows whose counts are all 0: SlimeCore's
// (SlimeCore.cpp:728-773, with a threads argument) keeps the center and the
// platform height, SlimeCoreGPU's (SlimeCoreGPU.cu:3958-3990) takes the first
// point and height in its order. A C caller's arity cannot be seen, so the
// GPU library's rule is the default and SLIME_REFINE_RULE=cpu in the
// environment (read at every call) selects SlimeCore's, for callers that load
// this library in place of the CPU one. The extensions that refine the same
// way (the many-world and the allocation-free forms) follow the same switch.
int32_t shared_refine_rule() {
    const char* v = std::getenv("SLIME_REFINE_RULE");
    return v && std::strcmp(v, "cpu") == 0 ? slime::kRuleKeepCenter : slime::kRuleFirstOrder;
}

constexpr int64_t kCpuRdLimit = (INT32_MAX - 17LL) / 2LL;
constexpr int64_t kGpuRdLimit = (INT32_MAX - 19LL) / 2LL;

}  // namespace

extern "C" {

int64_t search_slime_clusters_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                                       int32_t max_size, int64_t rd_min, int32_t search_center_x,
                                       int32_t search_center_z, ExtChunkResult* results_buffer,
                                       int32_t max_results, int32_t threads, int32_t precise_afk) {
    (void)threads;
    const int64_t total = search_common(seed, rd_max, min_size, max_size, rd_min, search_center_x,
                                        search_center_z, results_buffer, max_results, kCpuRdLimit);
    if (precise_afk != 0 && total > 0) precise_cpu(seed, results_buffer, total, max_results);
    return total;
}

int64_t search_slime_clusters(int64_t seed, int64_t rd_max, int32_t min_size, int32_t max_size,
                              int64_t rd_min, ExtChunkResult* results_buffer, int32_t max_results,
                              int32_t threads, int32_t precise_afk) {
    return search_slime_clusters_centered(seed, rd_max, min_size, max_size, rd_min, 0, 0,
                                          results_buffer, max_results, threads, precise_afk);
}

int64_t search_slime_clusters_gpu_v1_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                                              int32_t max_size, int64_t rd_min,
                                              int32_t search_center_x, int32_t search_center_z,
                                              ExtChunkResult* results_buffer, int32_t max_results,
                                              int32_t precise_afk) {
    const int64_t total = search_common(seed, rd_max, min_size, max_size, rd_min, search_center_x,
                                        search_center_z, results_buffer, max_results, kGpuRdLimit);
    if (precise_afk != 0 && total > 0) precise_gpu(seed, results_buffer, total, max_results);
    return total;
}

int64_t search_slime_clusters_gpu_v1(int64_t seed, int64_t rd_max, int32_t min_size,
                                     int32_t max_size, int64_t rd_min,
                                     ExtChunkResult* results_buffer, int32_t max_results,
                                     int32_t precise_afk) {
    return search_slime_clusters_gpu_v1_centered(seed, rd_max, min_size, max_size, rd_min, 0, 0,
                                                 results_buffer, max_results, precise_afk);
}

int64_t search_slime_clusters_gpu_v34(int64_t seed, int64_t rd_max, int32_t min_size,
                                      int32_t max_size, int64_t rd_min,
                                      ExtChunkResult* results_buffer, int32_t max_results,
                                      int32_t precise_afk) {
    return search_slime_clusters_gpu_v1(seed, rd_max, min_size, max_size, rd_min, results_buffer,
                                        max_results, precise_afk);
}

int64_t search_slime_clusters_gpu_v34_centered(int64_t seed, int64_t rd_max, int32_t min_size,
                                               int32_t max_size, int64_t rd_min,
                                               int32_t search_center_x, int32_t search_center_z,
                                               ExtChunkResult* results_buffer,
                                               int32_t max_results, int32_t precise_afk) {
    return search_slime_clusters_gpu_v1_centered(seed, rd_max, min_size, max_size, rd_min,
                                                 search_center_x, search_center_z,
                                                 results_buffer, max_results, precise_afk);
}

int64_t slime_scan_rect(int64_t seed, int32_t x0, int32_t z0, int64_t w, int64_t h,
                        int32_t min_size, int32_t max_size, ExtChunkResult* results_buffer,
                        int32_t max_results, int64_t* hist222) {
    if (!results_buffer || max_results <= 0 || w <= 0 || h <= 0 || w > INT32_MAX ||
        h > INT32_MAX || !slime::within_search_limit(w, h) || min_size < 0 || min_size > 221 ||
        max_size < min_size)
        return -1;
    slime::SearchSpec spec;
    spec.seed = seed;
    spec.x0 = x0;
    spec.z0 = z0;
    spec.w = w;
    spec.h = h;
    spec.min_size = min_size;
    spec.max_size = max_size;
    return slime::run_search(spec, reinterpret_cast<slime::ResultRow*>(results_buffer),
                             max_results, hist222);
}

int64_t slime_scan_bands(int64_t seed, int32_t x0, int32_t z0, int64_t w, int64_t h,
                         int32_t band_rows, int32_t mode, int32_t min_size, int32_t max_size,
                         ExtChunkResult* results_buffer, int32_t max_results, int64_t* hist222) {
    if (!results_buffer || max_results <= 0 || w <= 0 || h <= 0 || w > INT32_MAX ||
        h > INT32_MAX || !slime::within_search_limit(w, h) || band_rows < 0 || mode < 0 ||
        mode > 5 || min_size < 0 || min_size > 221 || max_size < min_size)
        return -1;
    slime::SearchSpec spec;
    spec.seed = seed;
    spec.x0 = x0;
    spec.z0 = z0;
    spec.w = w;
    spec.h = h;
    spec.min_size = min_size;
    spec.max_size = max_size;
    spec.band_rows = band_rows;
    spec.mode = (slime::ScanMode)mode;
    return slime::run_search(spec, reinterpret_cast<slime::ResultRow*>(results_buffer),
                             max_results, hist222);
}

void slime_kernel_stats(int32_t kind, int64_t* launches, int64_t* kernel_ns, int64_t* centers,
                        int32_t reset) {
    slime::kernel_stats(kind, launches, kernel_ns, centers, reset != 0);
}

int32_t slime_device_info(int32_t* compute_units, int32_t* clock_khz, int64_t* hbm_bytes) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return 0;
}

void request_cancel(void) {
    control().cancel.store(1, std::memory_order_relaxed);
    control().pause.store(0, std::memory_order_relaxed);
}

void request_pause(void) { control().pause.store(1, std::memory_order_relaxed); }

void resume_search(void) { control().pause.store(0, std::memory_order_relaxed); }

void reset_cancel(void) {
    control().cancel.store(0, std::memory_order_relaxed);
    control().pause.store(0, std::memory_order_relaxed);
    control().progress.store(0, std::memory_order_relaxed);
    control().processed_centers.store(0, std::memory_order_relaxed);
    control().scan_work_ns.store(0, std::memory_order_relaxed);
}

int32_t get_progress(void) { return control().progress.load(std::memory_order_relaxed); }

int64_t get_processed_centers(void) {
    return control().processed_centers.load(std::memory_order_relaxed);
}

int64_t get_gpu_scan_work_ns(void) { return control().scan_work_ns.load(std::memory_order_relaxed); }

void set_y_scan_config(int32_t enabled, int32_t platform_y, int32_t y_min, int32_t y_max,
                       int32_t y_step) {
    g_yscan.enabled = enabled ? 1 : 0;
    g_yscan.platform_y = platform_y;
    g_yscan.y_min = y_min;
    g_yscan.y_max = y_max;
    g_yscan.y_step = y_step > 0 ? y_step : 4;
}

void score_candidates_precise(int64_t seed, ExtChunkResult* results_buffer, int32_t count,
                              int32_t threads, int32_t scan_xz) {
    (void)threads;
    if (!results_buffer || count <= 0) return;
    control().progress.store(90, std::memory_order_relaxed);
    if (cancelled()) return;
    (void)slime::refine_rows(
        refine_spec(seed, scan_xz ? 1 : 0, slime::kRuleKeepCenter, g_yscan.enabled),
        reinterpret_cast<slime::ResultRow*>(results_buffer), count);
    control().progress.store(99, std::memory_order_relaxed);
}

void refine_candidates_y(int64_t seed, ExtChunkResult* results_buffer, int32_t count,
                         int32_t threads) {
    (void)threads;
    if (!results_buffer || count <= 0) return;
    control().progress.store(90, std::memory_order_relaxed);
    (void)slime::refine_rows(refine_spec(seed, 1, shared_refine_rule(), 1),
                             reinterpret_cast<slime::ResultRow*>(results_buffer), count);
    control().progress.store(99, std::memory_order_relaxed);
}

int32_t refine_candidates_y_noalloc(int64_t seed, ExtChunkResult* results_buffer,
                                    int32_t count) {
    if (count < 0 || (count > 0 && !results_buffer)) return -1;
    if (count > slime::kNoAllocMaxRows) return -2;
    if (count == 0) return 0;
    control().progress.store(90, std::memory_order_relaxed);
    const int rc = slime::refine_rows_noalloc(refine_spec(seed, 1, shared_refine_rule(), 1),
                                              reinterpret_cast<slime::ResultRow*>(results_buffer),
                                              count);
    control().progress.store(99, std::memory_order_relaxed);
    return rc;
}

int32_t refine_candidates_y_multi(int32_t nworlds, const int64_t* world_seeds, const int32_t* first,
                                  ExtChunkResult* results_buffer) {
    if (nworlds < 0 || (nworlds > 0 && (!world_seeds || !first || !results_buffer)) ||
        (nworlds > 0 && first[0] != 0))
        return -1;
    if (nworlds == 0) return 0;
    for (int32_t w = 0; w < nworlds; ++w)
        if (first[w + 1] < first[w]) return -1;
    const int32_t count = first[nworlds];
    if (count == 0) return 0;
    std::vector<int64_t> seeds((size_t)count);
    for (int32_t w = 0; w < nworlds; ++w)
        std::fill(seeds.begin() + first[w], seeds.begin() + first[w + 1], world_seeds[w]);
    return slime::refine_rows_multi(refine_spec(0, 1, shared_refine_rule(), 1), seeds.data(),
                                    reinterpret_cast<slime::ResultRow*>(results_buffer), count);
}

void slime_refine_stats(int64_t* kernel_ns, int64_t* evaluations) {
    slime::refine_last_stats(kernel_ns, evaluations);
}

bool is_slime_chunk_c(int64_t seed, int64_t cx, int64_t cz) {
    return slime::is_slime(seed, (int32_t)cx, (int32_t)cz);
}

int32_t get_cuda_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int32_t get_cuda_device_name(char* buffer, int32_t buffer_size) {
    if (!buffer || buffer_size <= 1) return 0;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        buffer[0] = '\0';
        return 0;
    }
    const size_t n = std::min((size_t)buffer_size - 1, std::strlen(prop.name));
    std::memcpy(buffer, prop.name, n);
    buffer[n] = '\0';
    return (int32_t)n;
}

void cleanup_gpu_resources(void) {
    slime::release_refine();
    slime::release_engine();
}

// The reference's shape numbering (SlimeCoreGPU.cu:31: 1 = 128x8, 2 = 256x4,
// 3 = 256x8, 4 = 512x4 threads x cells per thread; 0 before any scan): the
// plain kernel runs 256 threads x 4 columns, the banded kernel 256 threads x
// 8 chunks.
int32_t get_gpu_v1_shape(void) {
    const int k = slime::last_scan_kernel();
    return k < 0 ? 0 : k == 16 ? 2 : 3;
}
int32_t get_gpu_v1_algorithm(void) { return slime::last_scan_kernel(); }
int32_t get_gpu_v1_rng(void) { return g_rng_override >= 0 ? g_rng_override : 0; }
void set_gpu_v1_rng_override(int32_t value) {
    g_rng_override = (value >= 0 && value <= 2) ? value : -1;
}
int32_t get_gpu_v34_shape(void) { return get_gpu_v1_shape(); }
int32_t get_gpu_v34_rng(void) { return get_gpu_v1_rng(); }
void set_gpu_v34_rng_override(int32_t value) { set_gpu_v1_rng_override(value); }

int32_t slime_search_batch(int32_t n, const int64_t* seeds, const int64_t* rd_max,
                           const int32_t* min_size, const int32_t* max_size, const int64_t* rd_min,
                           const int32_t* center_x, const int32_t* center_z,
                           ExtChunkResult* results_buffer, int32_t max_results, int64_t* totals,
                           int32_t flags) {
    if (n < 0 || (n > 0 && (!seeds || !rd_max || !min_size || !max_size || !rd_min ||
                            !center_x || !center_z || !results_buffer || !totals)) ||
        max_results <= 0)
        return -1;
    std::vector<slime::SearchSpec> specs;
    std::vector<int32_t> index;
    for (int32_t i = 0; i < n; ++i) {
        totals[i] = -1;
        if (!valid_args(rd_max[i], kGpuRdLimit, min_size[i], max_size[i], rd_min[i], center_x[i],
                        center_z[i], results_buffer, max_results))
            continue;
        specs.push_back(square_spec(seeds[i], rd_max[i], min_size[i], max_size[i], rd_min[i],
                                    center_x[i], center_z[i]));
        index.push_back(i);
    }
    if (specs.empty()) return 0;
    std::vector<slime::ResultRow> rows(specs.size() * (size_t)max_results);
    std::vector<int64_t> tot(specs.size(), -1);
    slime::SweepOptions opt;
    opt.use_graph = (flags & 1) != 0;
    opt.persistent = (flags & 2) != 0;
    if (slime::run_batch(specs.data(), (int32_t)specs.size(), max_results, rows.data(),
                         tot.data(), opt, nullptr) != 0)
        return -1;
    for (size_t k = 0; k < specs.size(); ++k) {
        const int32_t i = index[k];
        totals[i] = tot[k];
        const size_t m = (size_t)std::max<int64_t>(0, std::min<int64_t>(tot[k], max_results));
        std::memcpy(results_buffer + (size_t)i * max_results, rows.data() + k * max_results,
                    sizeof(ExtChunkResult) * m);
    }
    return 0;
}

void slime_batch_stats(int64_t* centers, int64_t* kernel_ns, int64_t* prepare_ns,
                       int64_t* sweep_ns, int64_t* wall_ns, int32_t* launches, int32_t* jobs,
                       int32_t* fallbacks, int64_t* hits_stored, int64_t* hits_copied) {
    slime::BatchStats st;
    slime::last_batch_stats(&st);
    if (centers) *centers = st.centers;
    if (kernel_ns) *kernel_ns = st.kernel_ns;
    if (prepare_ns) *prepare_ns = st.prepare_ns;
    if (sweep_ns) *sweep_ns = st.sweep_ns;
    if (wall_ns) *wall_ns = st.wall_ns;
    if (launches) *launches = st.launches;
    if (jobs) *jobs = st.jobs;
    if (fallbacks) *fallbacks = st.fallbacks;
    if (hits_stored) *hits_stored = st.hits_stored;
    if (hits_copied) *hits_copied = st.hits_copied;
}

int32_t slime_batch_class_centers(int64_t* centers, int32_t cap) {
    if (!centers || cap < 0) return -1;
    slime::BatchStats st;
    slime::last_batch_stats(&st);
    const int32_t n = std::min<int32_t>(cap, slime::kSweepClasses);
    for (int32_t c = 0; c < n; ++c) centers[c] = st.class_centers[c];
    return n;
}

int32_t slime_sweep_costs(double* ns_row, double* extra_rows, int32_t* nb, int32_t* ph,
                          int32_t cap) {
    double a[slime::kSweepClasses], b[slime::kSweepClasses];
    if (!slime::sweep_cost_table(a, b)) {
        (void)hipGetLastError();
        return -1;
    }
    const int32_t n = std::min<int32_t>(cap, slime::kSweepClasses);
    for (int32_t c = 0; c < n; ++c) {
        if (ns_row) ns_row[c] = a[c];
        if (extra_rows) extra_rows[c] = b[c];
        if (nb) nb[c] = slime::kSweepClass[c].nb;
        if (ph) ph[c] = slime::kSweepClass[c].ph;
    }
    return n;
}

void* slime_session_create_ex(int64_t hbm_budget, int32_t max_active, int32_t round_ms,
                              int32_t flags) {
    if (hbm_budget <= 0 || max_active <= 0 || round_ms <= 0 || (flags & ~1)) return nullptr;
    auto* s = new (std::nothrow) slime::Session(hbm_budget, max_active, round_ms, (flags & 1) != 0);
    if (!s) return nullptr;
    if (!s->ok()) {
        delete s;
        (void)hipGetLastError();
        return nullptr;
    }
    return s;
}

void* slime_session_create(int64_t hbm_budget, int32_t max_active, int32_t round_ms) {
    return slime_session_create_ex(hbm_budget, max_active, round_ms, 0);
}

void slime_session_destroy(void* session) { delete static_cast<slime::Session*>(session); }

int64_t slime_session_submit(void* session, int64_t seed, int64_t rd_max, int32_t min_size,
                             int32_t max_size, int64_t rd_min, int32_t search_center_x,
                             int32_t search_center_z, int32_t max_results) {
    ExtChunkResult probe;  // valid_args checks the buffer pointer only
    if (!session || !valid_args(rd_max, kGpuRdLimit, min_size, max_size, rd_min,
                                search_center_x, search_center_z, &probe, max_results))
        return -1;
    return static_cast<slime::Session*>(session)->submit(
        square_spec(seed, rd_max, min_size, max_size, rd_min, search_center_x, search_center_z),
        max_results);
}

int32_t slime_session_pause(void* session, int64_t id) {
    return session && static_cast<slime::Session*>(session)->pause(id) ? 0 : -1;
}

int32_t slime_session_resume(void* session, int64_t id) {
    return session && static_cast<slime::Session*>(session)->resume(id) ? 0 : -1;
}

int32_t slime_session_cancel(void* session, int64_t id) {
    return session && static_cast<slime::Session*>(session)->cancel(id) ? 0 : -1;
}

int32_t slime_session_poll(void* session, int64_t id, int32_t* progress_pct, int64_t* total,
                           int64_t* latency_ns) {
    if (!session) return -1;
    slime::ReqState st;
    double prog = 0;
    if (!static_cast<slime::Session*>(session)->poll(id, &st, &prog, total, latency_ns)) return -1;
    if (progress_pct) *progress_pct = (int32_t)(prog * 100.0);
    return (int32_t)st;
}

int32_t slime_session_wait(void* session, int64_t id, int32_t timeout_ms) {
    if (!session) return -1;
    return (int32_t) static_cast<slime::Session*>(session)->wait(id, timeout_ms);
}

int64_t slime_session_result(void* session, int64_t id, ExtChunkResult* results_buffer,
                             int32_t cap) {
    if (!session) return -1;
    return static_cast<slime::Session*>(session)->result(
        id, reinterpret_cast<slime::ResultRow*>(results_buffer), cap);
}

int32_t slime_session_release(void* session, int64_t id) {
    return session && static_cast<slime::Session*>(session)->release(id) ? 0 : -1;
}

void slime_session_stats(void* session, int64_t* centers_done, int64_t* kernel_ns,
                         int32_t* rounds, int32_t* active, int32_t* queued, int64_t* hbm_in_use,
                         int64_t* hbm_budget) {
    if (!session) return;
    const slime::SessionStats st = static_cast<slime::Session*>(session)->stats();
    if (centers_done) *centers_done = st.centers_done;
    if (kernel_ns) *kernel_ns = st.kernel_ns;
    if (rounds) *rounds = st.rounds;
    if (active) *active = st.active;
    if (queued) *queued = st.queued;
    if (hbm_in_use) *hbm_in_use = st.hbm_in_use;
    if (hbm_budget) *hbm_budget = st.hbm_budget;
}

}  // extern "C"
