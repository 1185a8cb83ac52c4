// Host orchestration of one search: a list of kernel launches (banded
// first-draw slabs where the region is wide enough, plain slabs elsewhere),
// the score histogram, the emit gate, overflow retry, the exact correction of
// Java's rejected first draws, and the top-K ranking.
//
// Reference behaviour followed (SlimeCoreGPU.cu:3164-3915):
//   - the region is scanned in slabs; after each launch the hit counter and
//     buffer are read back; an overflowing launch is rerun (:3795-3805);
//   - every in-range hit is counted, but only hits that can still reach the
//     global top-K are materialised (emit gate, :3731-3740);
//   - ranking: size desc, center_x asc, center_z asc (:213-220, 3860-3880);
//   - pause is a sleep-poll between launches, cancel ranks what was found
//     (:3754-3757).
// The gate here is exact and needs no completion pass: it is raised only to a
// score s for which at least K hits of score >= s were counted by centers
// whose count is already final.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <new>
#include <thread>

#include "engine.h"
#include "search_util.h"
#include "slime_math.h"

namespace slime {

Control& control() {
    static Control c;
    return c;
}

namespace {

constexpr uint32_t kHitCap = 1u << 24;                       // hits per launch (256 MiB)
constexpr size_t kEagerHits = 4096;  // hits copied back with every launch's counters (64 KB)
constexpr int64_t kPlainCentersPerLaunch = int64_t(1) << 37;
// Centers per banded launch at most: each launch ends in a partial last wave
// of tiles, so fewer, longer launches waste less (full world: 2^39 9.90, 2^40
// 9.99-10.02, 2^41 10.06-10.07 T centers/s end to end); pause and cancel act
// between launches (about 0.2 s apart at this size).
constexpr int64_t kBandCentersPerLaunch = int64_t(1) << 41;
constexpr int kBandTileH = 2048;
// Small regions get short tiles so that their few tiles spread over the
// device: each tile's rows are generated one step after another, so a search
// of a few thousand rows is bound by one tile's steps (radius 2,048: 0.275 ->
// 0.237 ms per call with 16 rows instead of 64, for 1.6x the generated rows).
constexpr int kMinBandTileH = 16;
// Bands per banded launch at most (kernel instantiations 1..14, all at three
// or more waves per SIMD without spills); the SLIME_MAX_BANDS environment
// variable lowers it, like the reference's SLIME_GPU_V1_* path switches
// (SlimeCoreGPU.cu:3267-3354).
int g_max_bands = [] {
    const char* v = std::getenv("SLIME_MAX_BANDS");
    const int n = v ? std::atoi(v) : 0;
    return n >= 1 && n <= kMaxBands ? n : kMaxBands;
}();

// Host array of center fixes, in pinned memory where it can be had (the copies
// to and from the device then need no staging), kept across searches so that a
// full-world search's 2.3 M fixes (55 MB) find their pages already mapped.
class FixList {
  public:
    ~FixList() { release(); }
    void release() {
        if (data_) {
            if (pinned_) (void)hipHostFree(data_);
            else delete[] data_;
        }
        data_ = nullptr;
        n_ = cap_ = 0;
    }
    void clear() { n_ = 0; }
    // false only if no memory at all could be had
    bool reserve(size_t want) {
        if (want <= cap_) return true;
        CenterFix* p = nullptr;
        bool pinned = hipHostMalloc((void**)&p, sizeof(CenterFix) * want, hipHostMallocDefault) ==
                      hipSuccess;
        if (!pinned) {
            (void)hipGetLastError();
            p = new (std::nothrow) CenterFix[want];
            if (!p) return false;
        }
        if (n_) std::memcpy(p, data_, sizeof(CenterFix) * n_);
        const size_t n = n_;
        release();
        data_ = p;
        pinned_ = pinned;
        n_ = n;
        cap_ = want;
        return true;
    }
    void push_back(const CenterFix& f) {
        if (n_ == cap_ && !reserve(std::max<size_t>(1024, 2 * cap_))) {
            failed_ = true;
            return;
        }
        data_[n_++] = f;
    }
    bool failed() const { return failed_; }
    void reset_failed() { failed_ = false; }
    CenterFix* data() { return data_; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    CenterFix* begin() { return data_; }
    CenterFix* end() { return data_ + n_; }

  private:
    CenterFix* data_ = nullptr;
    size_t n_ = 0, cap_ = 0;
    bool pinned_ = false;
    bool failed_ = false;
};

struct Buffers {
    int device = -1;
    int cus = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    // the first-draw mismatch search and recount, while the first scan
    // launch runs on `stream` (default priority, like `stream`)
    hipStream_t aux = nullptr;
    unsigned int* h_mm_count = nullptr;  // pinned
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    Hit* d_hits = nullptr;
    unsigned long long* d_counters = nullptr;  // [0] emitted, [1..222] histogram
    unsigned long long* h_counters = nullptr;  // pinned
    Hit* h_hits = nullptr;                     // pinned
    size_t h_hits_cap = 0;
    uint64_t* d_table = nullptr;
    int32_t table_n = 0;
    MismatchChunk* d_mismatch = nullptr;
    unsigned int* d_mismatch_count = nullptr;
    CenterFix* d_fixes = nullptr;
    size_t fixes_cap = 0;
};
FixList g_fixes;  // the host fixes of the current search (guarded by g_mutex)

Buffers g_buf;
std::mutex g_mutex;
// Per kernel kind (0 plain, 1 banded): launches, device ns, centers decided.
// [plain, banded, banded with three x-phases (also counted as banded)][launches, ns, centers]
std::atomic<int64_t> g_kstat[3][3];

void free_buffers() {
    if (g_buf.device >= 0) (void)hipSetDevice(g_buf.device);
    if (g_buf.d_hits) (void)hipFree(g_buf.d_hits);
    if (g_buf.d_counters) (void)hipFree(g_buf.d_counters);
    if (g_buf.h_counters) (void)hipHostFree(g_buf.h_counters);
    if (g_buf.h_hits) (void)hipHostFree(g_buf.h_hits);
    if (g_buf.d_table) (void)hipFree(g_buf.d_table);
    if (g_buf.d_mismatch) (void)hipFree(g_buf.d_mismatch);
    if (g_buf.d_mismatch_count) (void)hipFree(g_buf.d_mismatch_count);
    if (g_buf.d_fixes) (void)hipFree(g_buf.d_fixes);
    if (g_buf.ev_start) (void)hipEventDestroy(g_buf.ev_start);
    if (g_buf.ev_stop) (void)hipEventDestroy(g_buf.ev_stop);
    if (g_buf.stream) (void)hipStreamDestroy(g_buf.stream);
    if (g_buf.aux) (void)hipStreamDestroy(g_buf.aux);
    if (g_buf.h_mm_count) (void)hipHostFree(g_buf.h_mm_count);
    g_fixes.release();
    g_buf = Buffers{};
}

bool ensure_buffers() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (g_buf.device == dev && g_buf.d_hits) return true;
    free_buffers();
    g_buf.device = dev;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
        g_buf.cus = cus;
    const std::vector<uint64_t>& table = mismatch_table();
    g_buf.table_n = (int32_t)table.size();
    if (hipStreamCreateWithFlags(&g_buf.stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&g_buf.aux, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc(&g_buf.h_mm_count, sizeof(unsigned int), hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreate(&g_buf.ev_start) != hipSuccess ||
        hipEventCreate(&g_buf.ev_stop) != hipSuccess ||
        hipMalloc(&g_buf.d_hits, sizeof(Hit) * (size_t)kHitCap) != hipSuccess ||
        hipMalloc(&g_buf.d_counters, sizeof(unsigned long long) * (1 + kHistBins)) != hipSuccess ||
        hipHostMalloc(&g_buf.h_counters, sizeof(unsigned long long) * (1 + kHistBins),
                      hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&g_buf.d_table, sizeof(uint64_t) * table.size()) != hipSuccess ||
        hipMemcpy(g_buf.d_table, table.data(), sizeof(uint64_t) * table.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&g_buf.d_mismatch, sizeof(MismatchChunk) * kMismatchCap) != hipSuccess ||
        hipMalloc(&g_buf.d_mismatch_count, sizeof(unsigned int)) != hipSuccess) {
        free_buffers();
        return false;
    }
    return true;
}

bool ensure_host_hits(size_t n) {
    if (g_buf.h_hits_cap >= n) return true;
    if (g_buf.h_hits) (void)hipHostFree(g_buf.h_hits);
    g_buf.h_hits = nullptr;
    g_buf.h_hits_cap = 0;
    const size_t want = std::max<size_t>(n, 1 << 16);
    if (hipHostMalloc(&g_buf.h_hits, sizeof(Hit) * want, hipHostMallocDefault) != hipSuccess) {
        g_buf.h_hits = nullptr;
        return false;
    }
    g_buf.h_hits_cap = want;
    return true;
}

bool ensure_fixes(size_t n) {
    if (g_buf.fixes_cap >= n) return true;
    if (g_buf.d_fixes) (void)hipFree(g_buf.d_fixes);
    g_buf.d_fixes = nullptr;
    g_buf.fixes_cap = 0;
    const size_t want = std::max<size_t>(n, 1 << 16);
    if (hipMalloc(&g_buf.d_fixes, sizeof(CenterFix) * want) != hipSuccess) {
        g_buf.d_fixes = nullptr;
        return false;
    }
    g_buf.fixes_cap = want;
    return true;
}

// Score bookkeeping shared by every launch of one search.
struct Ranker {
    int32_t min_size = 0, max_size = 221;
    size_t keep = 0;
    int64_t hist[kHistBins] = {0};
    int64_t pending[kHistBins] = {0};  // first-draw scores of centers to be recounted
    std::vector<Hit> cand;
    const MismatchIndex* recount = nullptr;  // centers whose scan count is provisional
    int32_t emit_min = 0;

    int32_t gate(const unsigned long long* extra) const {
        int64_t h[kHistBins];
        for (int s = 0; s < kHistBins; ++s)
            h[s] = hist[s] - pending[s] + (extra ? (int64_t)extra[s] : 0);
        return gate_for(h, (int64_t)keep, min_size);
    }
    void add(const unsigned long long* slab_hist, const Hit* hits, size_t n) {
        for (int s = 0; s < kHistBins; ++s) hist[s] += (int64_t)slab_hist[s];
        for (size_t i = 0; i < n; ++i)
            if (!recount || !recount->affects(hits[i].cx, hits[i].cz)) cand.push_back(hits[i]);
        if (cand.size() > std::max<size_t>(2 * keep, 1 << 20)) {
            std::nth_element(cand.begin(), cand.begin() + keep, cand.end(), ranks_before);
            cand.resize(keep);
        }
        emit_min = std::max(emit_min, gate(nullptr));
    }
};

// One kernel launch of the plan.
struct Work {
    bool banded = false;
    ScanParams plain{};
    BandParams band{};
    int nb = 0, ph = 2;
    bool union_bound = false;
    int64_t centers = 0;  // candidate centers this launch decides
};

enum class RunResult { kOk, kSplit, kError };

// while_running (if set) runs once, after the first launch is queued and
// before the host waits for it (host work that the launch hides).
RunResult run_work(Work& wk, Ranker& rk, const std::function<bool()>& while_running = {}) {
    bool ran = false;
    // The first hits come back with the counters (no second round trip when a
    // launch emits at most this many); the rest are copied after the count is
    // known.
    const size_t eager = std::min<size_t>(kEagerHits, (size_t)kHitCap);
    if (!ensure_host_hits(eager)) return RunResult::kError;
    for (;;) {
        hipError_t launch_err;
        // a fill kernel: cheaper to issue than hipMemsetAsync
        if (launch_fill_u32(reinterpret_cast<uint32_t*>(g_buf.d_counters), 2 * (1 + kHistBins),
                            0u, g_buf.stream) != hipSuccess ||
            hipEventRecord(g_buf.ev_start, g_buf.stream) != hipSuccess)
            return RunResult::kError;
        if (wk.banded) {
            BandParams& p = wk.band;
            p.emit_min = rk.emit_min;
            p.hits = g_buf.d_hits;
            p.hit_cap = kHitCap;
            p.emitted = g_buf.d_counters;
            p.hist = g_buf.d_counters + 1;
            launch_err = launch_scan_band(p, wk.nb, wk.ph, wk.union_bound, g_buf.stream);
        } else {
            ScanParams& p = wk.plain;
            p.emit_min = rk.emit_min;
            p.hits = g_buf.d_hits;
            p.hit_cap = kHitCap;
            p.emitted = g_buf.d_counters;
            p.hist = g_buf.d_counters + 1;
            launch_err = launch_scan_plain(p, g_buf.stream);
        }
        if (launch_err != hipSuccess || hipEventRecord(g_buf.ev_stop, g_buf.stream) != hipSuccess ||
            hipMemcpyAsync(g_buf.h_counters, g_buf.d_counters,
                           sizeof(unsigned long long) * (1 + kHistBins), hipMemcpyDeviceToHost,
                           g_buf.stream) != hipSuccess ||
            hipMemcpyAsync(g_buf.h_hits, g_buf.d_hits, sizeof(Hit) * eager,
                           hipMemcpyDeviceToHost, g_buf.stream) != hipSuccess)
            return RunResult::kError;
        if (while_running && !ran) {
            ran = true;
            if (!while_running()) return RunResult::kError;
        }
        if (hipStreamSynchronize(g_buf.stream) != hipSuccess) return RunResult::kError;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, g_buf.ev_start, g_buf.ev_stop) == hipSuccess) {
            std::atomic<int64_t>* st = g_kstat[wk.banded ? 1 : 0];
            st[0].fetch_add(1, std::memory_order_relaxed);
            st[1].fetch_add((int64_t)((double)ms * 1e6), std::memory_order_relaxed);
            st[2].fetch_add(wk.centers, std::memory_order_relaxed);
            if (wk.banded && wk.ph == 3) {
                g_kstat[2][0].fetch_add(1, std::memory_order_relaxed);
                g_kstat[2][1].fetch_add((int64_t)((double)ms * 1e6), std::memory_order_relaxed);
                g_kstat[2][2].fetch_add(wk.centers, std::memory_order_relaxed);
            }
        }
        const unsigned long long emitted = g_buf.h_counters[0];
        const unsigned long long* slab_hist = g_buf.h_counters + 1;
        if (emitted > kHitCap) {
            const int32_t next = rk.gate(slab_hist);
            if (next > rk.emit_min) {
                rk.emit_min = next;
                continue;
            }
            return RunResult::kSplit;
        }
        if (emitted > eager) {
            if (!ensure_host_hits(emitted) ||
                hipMemcpyAsync(g_buf.h_hits, g_buf.d_hits, sizeof(Hit) * emitted,
                               hipMemcpyDeviceToHost, g_buf.stream) != hipSuccess ||
                hipStreamSynchronize(g_buf.stream) != hipSuccess)
                return RunResult::kError;
        }
        rk.add(slab_hist, g_buf.h_hits, emitted);
        return RunResult::kOk;
    }
}

// Halves a work along z; false if it cannot be split further.
bool split_work(const Work& wk, Work& a, Work& b) {
    a = b = wk;
    if (wk.banded) {
        const BandParams& p = wk.band;
        const int32_t rows = p.s_end - p.s_begin;
        if (rows <= p.tile_h) return false;
        const int32_t half = (rows / 2 + p.tile_h - 1) / p.tile_h * p.tile_h;
        a.band.s_end = p.s_begin + half;
        b.band.s_begin = p.s_begin + half;
        const int32_t tiles_x = p.tiles_x;
        a.band.tiles_total = tiles_x * ((a.band.s_end - a.band.s_begin + p.tile_h - 1) / p.tile_h);
        b.band.tiles_total = tiles_x * ((b.band.s_end - b.band.s_begin + p.tile_h - 1) / p.tile_h);
        a.centers = wk.centers * half / rows;
        b.centers = wk.centers - a.centers;
        return true;
    }
    const ScanParams& p = wk.plain;
    if (p.h <= 4) return false;
    const int32_t half = (p.h / 2 + 3) / 4 * 4;
    const int32_t th = std::min(p.tile_h, half);
    a.plain.h = half;
    a.plain.tile_h = th;
    a.plain.tiles_total = p.tiles_x * ((half + th - 1) / th);
    b.plain.z0 = (int32_t)((uint32_t)p.z0 + (uint32_t)half);
    b.plain.h = p.h - half;
    b.plain.tile_h = th;
    b.plain.tiles_total = p.tiles_x * ((b.plain.h + th - 1) / th);
    a.centers = (int64_t)p.w * half;
    b.centers = (int64_t)p.w * b.plain.h;
    return true;
}

void fill_plain_filters(const SearchSpec& spec, int32_t& min_size, int32_t& max_size,
                  int32_t& use_annulus, int64_t& rd_min_sq, int32_t& acx, int32_t& acz) {
    min_size = spec.min_size;
    max_size = spec.max_size;
    use_annulus = spec.rd_min > 0;
    rd_min_sq = spec.rd_min * spec.rd_min;
    acx = spec.acx;
    acz = spec.acz;
}

// Plain slabs over centers [x0, x0+w) x [z0, z0+h).
void plan_plain(const SearchSpec& spec, int32_t x0, int64_t w, int32_t z0, int64_t h,
                std::deque<Work>& out) {
    const int32_t tiles_x = (int32_t)((w + kTileOut - 1) / kTileOut);
    const int64_t want_h = (h * tiles_x + 2047) / 2048;  // >= 2048 workgroups per launch
    const int32_t th = (int32_t)std::min<int64_t>(2048, std::max<int64_t>(16, (want_h + 3) / 4 * 4));
    int64_t slab = kPlainCentersPerLaunch / std::max<int64_t>(1, w);
    slab = std::max<int64_t>(th, slab / th * th);
    slab = std::min<int64_t>(slab, std::max<int64_t>(1, ((int64_t)1 << 30) / tiles_x) * th);
    for (int64_t z_off = 0; z_off < h; z_off += slab) {
        Work wk;
        ScanParams& p = wk.plain;
        const int64_t rows = std::min<int64_t>(slab, h - z_off);
        p.seed = spec.seed;
        p.x0 = x0;
        p.z0 = (int32_t)((uint32_t)z0 + (uint32_t)z_off);
        p.w = (int32_t)w;
        p.h = (int32_t)rows;
        p.tile_h = th;
        p.tiles_x = tiles_x;
        p.tiles_total = (int32_t)(tiles_x * ((rows + th - 1) / th));
        fill_plain_filters(spec, p.min_size, p.max_size, p.use_annulus, p.rd_min_sq, p.acx, p.acz);
        wk.centers = rows * w;
        out.push_back(wk);
    }
}

// Plain slabs over every row range of the spec (all rows, or each band).
void plan_plain_rows(const SearchSpec& spec, int32_t x0, int64_t w, std::deque<Work>& out) {
    if (spec.band_rows <= 0) {
        plan_plain(spec, x0, w, spec.z0, spec.h, out);
        return;
    }
    for (int64_t b = 0; b * kPeriod < spec.h; ++b) {
        const int64_t rows = std::min<int64_t>(spec.band_rows, spec.h - b * kPeriod);
        plan_plain(spec, x0, w, (int32_t)((uint32_t)spec.z0 + (uint32_t)(b * kPeriod)), rows, out);
    }
}

// ---- autotuner: measured cost of the banded kernels ------------------------
// A large search's band grouping, and whether its last, partial band is
// scanned as a short group of its own, are chosen by measured cost, as the
// reference times its kernel shapes on a saturated sample before a large
// search (SlimeCoreGPU.cu:3517-3599). Per (x-phases, bands) of the
// union-bound kernels, saturated launches of synthetic tiles at two tile
// heights give the cost of a tile as T * (rows + 16 + o): T ns per generated
// row, o fixed rows. Measured once per device at the first search of at
// least kTuneMinCenters centers; smaller searches, and SLIME_BAND_TUNE=0 in
// the environment, keep the even grouping.
constexpr int64_t kTuneMinCenters = int64_t(1) << 37;
constexpr int32_t kTileHaloRows = 16;  // generated rows beyond a tile's slot rows

struct BandCost {
    double T = 0, o = 0;
};
struct BandCosts {
    int device = -1;
    bool ready = false, failed = false;
    BandCost c[4][kUnionBoundMaxBands2 + 1];  // [x-phases][bands]
};
BandCosts g_costs;

bool band_tune_enabled() {
    const char* v = std::getenv("SLIME_BAND_TUNE");
    return !(v && std::strcmp(v, "0") == 0);
}

// SLIME_THREE_PHASE=0: the measured plan keeps the two-phase layout.
bool three_phase_enabled() {
    const char* v = std::getenv("SLIME_THREE_PHASE");
    return !(v && std::strcmp(v, "0") == 0);
}

// ns of one launch of nb bands and ph phases over `tiles` synthetic tiles of
// tile_h slot rows (nothing reaches min_size, so nothing is emitted).
bool time_band_launch(int nb, int ph, int32_t tile_h, int64_t want_tiles, double& ns_per_tile) {
    SearchSpec spec;
    spec.seed = 0x5EED5EED5EEDll + 17 * nb + ph;
    spec.x0 = -kPeriod;
    spec.z0 = 12345;
    spec.w = ph * kPeriod;
    const int64_t row_tiles = std::max<int64_t>(1, (want_tiles + kTilesPerPhase - 1) / kTilesPerPhase);
    const int64_t slot_rows = std::min<int64_t>(kPeriod / tile_h * tile_h, row_tiles * tile_h);
    spec.h = (int64_t)(nb - 1) * kPeriod + slot_rows;
    spec.band_rows = (int32_t)slot_rows;
    spec.min_size = kDiskCells;
    spec.max_size = kDiskCells;
    std::vector<BandJob> jobs;
    band_jobs(spec, x_parts(spec, ph), std::vector<int>{nb}, tile_h, int64_t(1) << 62, jobs);
    if (jobs.size() != 1) return false;
    BandParams p = jobs[0].p;
    p.emit_min = kDiskCells + 1;
    p.hits = g_buf.d_hits;
    p.hit_cap = 0;
    p.emitted = g_buf.d_counters;
    p.hist = g_buf.d_counters + 1;
    float ms = 0.f;
    if (hipEventRecord(g_buf.ev_start, g_buf.stream) != hipSuccess ||
        launch_scan_band(p, nb, ph, true, g_buf.stream) != hipSuccess ||
        hipEventRecord(g_buf.ev_stop, g_buf.stream) != hipSuccess ||
        hipStreamSynchronize(g_buf.stream) != hipSuccess ||
        hipEventElapsedTime(&ms, g_buf.ev_start, g_buf.ev_stop) != hipSuccess || p.tiles_total <= 0)
        return false;
    ns_per_tile = (double)ms * 1e6 / (double)p.tiles_total;
    return true;
}

bool calibrate_bands() {
    if (g_costs.device == g_buf.device && (g_costs.ready || g_costs.failed)) return g_costs.ready;
    g_costs = BandCosts{};
    g_costs.device = g_buf.device;
    g_costs.failed = true;
    const int64_t slots = 3 * (int64_t)std::max(1, g_buf.cus);  // workgroups resident at 3 per CU
    for (int ph = 1; ph <= 3; ++ph)
        for (int nb = min_bands(ph); nb <= union_bound_max_bands(ph); ++nb) {
            double warm = 0, a[2] = {0, 0};
            const int32_t ths[2] = {64, 1024};
            if (!time_band_launch(nb, ph, ths[0], slots, warm) ||
                !time_band_launch(nb, ph, ths[0], 24 * slots, a[0]) ||
                !time_band_launch(nb, ph, ths[1], 6 * slots, a[1]))
                return false;
            const double r0 = ths[0] + kTileHaloRows, r1 = ths[1] + kTileHaloRows;
            double T = (a[1] - a[0]) / (r1 - r0);
            if (!(T > 0)) T = a[1] / r1;
            g_costs.c[ph][nb].T = T;
            g_costs.c[ph][nb].o = std::max(0.0, a[0] / T - r0);
        }
    g_costs.failed = false;
    g_costs.ready = true;
    return true;
}

// Tile height: the largest (up to kBandTileH) that still gives every CU a few
// tiles when `col_tiles` x-tiles share each slot row; small regions trade halo
// rows for occupancy.
int32_t band_tile_h(int32_t slot_rows, int64_t col_tiles) {
    const int64_t want_tiles = 4 * (int64_t)g_buf.cus * 3;
    const int64_t fit_h = (slot_rows * col_tiles + want_tiles - 1) / want_tiles;
    return (int32_t)std::min<int64_t>(std::min<int64_t>(kBandTileH, (slot_rows + 3) / 4 * 4),
                                      std::max<int64_t>(kMinBandTileH, (fit_h + 3) / 4 * 4));
}

// Measured ns of a group of nb bands over s_rows slot rows of every x-part
// (negative where a part has no kernel of nb bands).
double group_ns(const std::vector<XPart>& parts, int nb, int32_t s_rows, int32_t tile_h) {
    double t = 0;
    for (const XPart& xp : parts) {
        if (nb < min_bands(xp.ph)) return -1.0;
        const BandCost& c = g_costs.c[xp.ph][nb];
        const int64_t tiles =
            (int64_t)xp.nblocks * xp.tiles_per_block() * ((s_rows + tile_h - 1) / tile_h);
        t += (double)tiles * c.T * ((double)tile_h + kTileHaloRows + c.o);
    }
    return t;
}

// Cheapest split of n consecutive bands into groups of at most max_bands
// (sizes largest first), each over s_rows slot rows; its measured ns, or -1
// when no split exists (three phases have no 1-band kernel).
double cheapest_groups(const std::vector<XPart>& parts, int64_t n, int max_bands, int32_t s_rows,
                       int32_t tile_h, std::vector<int>& sizes) {
    std::vector<double> cost((size_t)n + 1, -1.0);
    std::vector<int> pick((size_t)n + 1, 0);
    cost[0] = 0;
    for (int64_t k = 1; k <= n; ++k)
        for (int nb = 1; nb <= max_bands && nb <= k; ++nb) {
            const double g = group_ns(parts, nb, s_rows, tile_h);
            if (g < 0 || cost[k - nb] < 0) continue;
            const double v = cost[k - nb] + g;
            if (cost[k] < 0 || v < cost[k]) {
                cost[k] = v;
                pick[k] = nb;
            }
        }
    sizes.clear();
    if (cost[n] < 0) return -1.0;
    for (int64_t k = n; k > 0; k -= pick[k]) sizes.push_back(pick[k]);
    std::sort(sizes.begin(), sizes.end(), std::greater<int>());
    return cost[n];
}

// The tuned plan of one x-part (section (a)/(b_m) below), its jobs appended to
// `jobs`; its measured ns, or -1 when the part has no such plan.
double plan_part_tuned(const SearchSpec& spec, const XPart& xp, int max_bands, int32_t tile_h,
                       int64_t col_tiles, std::vector<BandJob>& jobs) {
    const std::vector<XPart> parts{xp};
    const int32_t slot_rows = slot_rows_of(spec);
    const int64_t nbz = bands_of(spec);
    // (a) every band over slot_rows in the cheapest groups. With a last,
    // partial band of rem rows after nf full ones, for m = 0..nf: (b_m) the
    // first nf - m bands over all P rows; the last m full bands and the
    // partial one over rows [0, rem); those m bands alone over rows [rem, P)
    // (m = 0: the partial band as a short group of its own). No launch then
    // scans the partial band's missing rows, and every group can keep a high
    // band count.
    std::vector<int> ga;
    const double ca = cheapest_groups(parts, nbz, max_bands, slot_rows, tile_h, ga);
    const int64_t nf = spec.h / kPeriod, rem = spec.h % kPeriod;
    double best = ca;
    int64_t best_m = -1;
    std::vector<int> g_head, g_low, g_high;
    int32_t th_low = 0, th_high = 0;
    if (spec.band_rows == 0 && nf >= 1 && rem > 0) {
        const int32_t thl = band_tile_h((int32_t)rem, col_tiles);
        const int32_t thh = band_tile_h((int32_t)(kPeriod - rem), col_tiles);
        for (int64_t m = 0; m <= nf; ++m) {
            std::vector<int> h, l, u;
            const double c0 = cheapest_groups(parts, nf - m, max_bands, kPeriod, tile_h, h);
            const double c1 = cheapest_groups(parts, m + 1, max_bands, (int32_t)rem, thl, l);
            const double c2 =
                m > 0 ? cheapest_groups(parts, m, max_bands, (int32_t)(kPeriod - rem), thh, u) : 0;
            if (c0 < 0 || c1 < 0 || c2 < 0) continue;
            const double c = c0 + c1 + c2;
            if (best < 0 || c < best) {
                best = c;
                best_m = m;
                g_head = h;
                g_low = l;
                g_high = u;
                th_low = thl;
                th_high = thh;
            }
        }
    }
    if (best < 0) return -1.0;
    if (best_m >= 0) {
        auto emit = [&](const std::vector<int>& sizes, int64_t b_first, int32_t s_lo, int32_t s_hi,
                        int32_t th) {
            for (const int nb : sizes) {
                band_jobs_part(spec, xp, b_first, nb, s_hi, th, kBandCentersPerLaunch, jobs, s_lo);
                b_first += nb;
            }
        };
        const int64_t b_tail = nf - best_m;
        emit(g_head, 0, 0, kPeriod, tile_h);
        emit(g_low, b_tail, 0, (int32_t)rem, th_low);
        emit(g_high, b_tail, (int32_t)rem, kPeriod, th_high);
    } else {
        band_jobs(spec, parts, ga, tile_h, kBandCentersPerLaunch, jobs);
    }
    return best;
}

// Banded launches over every center of the spec: two x-phases above P
// columns, one below (x_parts); bands grouped evenly by at most g_max_bands
// per launch, and at most union_bound_max_bands(ph) where the union bound
// applies (its kernels exist up to that many bands; at thresholds near 50 it
// beats the square bound on more bands). Where the measured costs are known
// (the first search of at least kTuneMinCenters centers on a device measures
// them), each x-part takes the cheapest grouping of the cost model instead,
// and a region of at most 5 bands, wider than 2P and not a whole number of 2P
// blocks wide, may take the three-phase layout (x_parts_three: no masked
// columns) when its measured cost is lower (SLIME_THREE_PHASE=0 keeps two
// phases). kBandedUnion3 forces that layout (tests).
void plan_banded(const SearchSpec& spec, std::deque<Work>& out) {
    const int ph = spec.w > kPeriod ? 2 : 1;
    const bool union_bound = union_bound_for(spec.mode, spec.min_size);
    const int64_t nbz = bands_of(spec);
    const bool force3 = spec.mode == ScanMode::kBandedUnion3 && spec.w > 2 * kPeriod &&
                        nbz >= kThreePhaseMinBands;
    auto max_bands_of = [&](const std::vector<XPart>& parts) {
        int mb = g_max_bands;
        if (union_bound)
            for (const XPart& xp : parts) mb = std::min(mb, union_bound_max_bands(xp.ph));
        return mb;
    };
    auto col_tiles_of = [](const std::vector<XPart>& parts) {
        int64_t t = 0;  // x-tiles of one slot row over all parts
        for (const XPart& xp : parts) t += (int64_t)xp.nblocks * xp.tiles_per_block();
        return t;
    };
    const int32_t slot_rows = slot_rows_of(spec);
    std::vector<BandJob> jobs;
    const bool tune = union_bound && !force3 && band_tune_enabled() &&
                      ((g_costs.device == g_buf.device && g_costs.ready) ||
                       (spec.w * spec.h >= kTuneMinCenters && calibrate_bands()));
    bool planned = false;
    if (tune) {
        std::vector<std::vector<XPart>> layouts{x_parts(spec, ph)};
        // the three-phase layout for regions of at most 5 bands: with more, the
        // two-phase parts keep wide band groups and the model's gain for the
        // three-phase part (at most 5 bands a group) did not hold (full world:
        // 9.86 vs 9.88 T centers/s of kernels)
        if (spec.w > 2 * kPeriod && spec.w % (2 * kPeriod) != 0 && three_phase_enabled() &&
            nbz <= kThreePhaseMaxBands)
            layouts.push_back(x_parts(spec, 3));
        // two phases where both lie inside the region, one phase (bands
        // paired) over the remaining columns
        if (spec.w > kPeriod && spec.w % (2 * kPeriod) != 0)
            layouts.push_back(x_parts_one_phase_tail(spec));
        double best = -1;
        for (const std::vector<XPart>& parts : layouts) {
            const int64_t col_tiles = col_tiles_of(parts);
            const int64_t ngroups = (nbz + max_bands_of(parts) - 1) / max_bands_of(parts);
            const int32_t tile_h = band_tile_h(slot_rows, col_tiles * ngroups);
            std::vector<BandJob> lj;
            double total = 0;
            for (const XPart& xp : parts) {
                const int mb = std::min(g_max_bands, union_bound_max_bands(xp.ph));
                const double c = plan_part_tuned(spec, xp, mb, tile_h, col_tiles, lj);
                if (c < 0) {
                    total = -1;
                    break;
                }
                total += c;
            }
            if (total >= 0 && (best < 0 || total < best)) {
                best = total;
                jobs.swap(lj);
            }
        }
        planned = best >= 0;
    }
    if (!planned) {
        jobs.clear();
        const std::vector<XPart> parts = x_parts(spec, force3 ? 3 : ph);
        const int max_bands = max_bands_of(parts);
        const int64_t ngroups = (nbz + max_bands - 1) / max_bands;
        std::vector<int> groups;
        for (int64_t gi = 0, left = nbz; gi < ngroups; ++gi) {
            const int64_t nb = (left + (ngroups - gi) - 1) / (ngroups - gi);
            groups.push_back((int)nb);
            left -= nb;
        }
        const int32_t tile_h = band_tile_h(slot_rows, col_tiles_of(parts) * ngroups);
        band_jobs(spec, parts, groups, tile_h, kBandCentersPerLaunch, jobs);
    }
    for (const BandJob& j : jobs) {
        Work wk;
        wk.banded = true;
        wk.nb = j.nb;
        wk.ph = j.ph;
        wk.union_bound = union_bound;
        wk.band = j.p;
        wk.band.tag = 0;
        wk.centers = j.centers;
        out.push_back(wk);
    }
}

// Finds the first-draw mismatches of the region and recounts every affected
// center exactly, in two halves on the aux stream: start_fixes queues the
// mismatch search, finish_fixes waits for it, collects the affected centers
// on the host, recounts them on the device and waits. The scan launches do
// not need them (they count first draws), so the first one runs meanwhile.
// Returns false on a device error.
bool start_fixes(const SearchSpec& spec) {
    return launch_fill_u32(g_buf.d_mismatch_count, 1, 0u, g_buf.aux) == hipSuccess &&
           launch_find_mismatches(spec.seed, (int32_t)((int64_t)spec.x0 - 8), spec.w + 16,
                                  (int32_t)((int64_t)spec.z0 - 8), spec.h + 16, g_buf.d_table,
                                  g_buf.table_n, g_buf.d_mismatch, kMismatchCap,
                                  g_buf.d_mismatch_count, g_buf.aux) == hipSuccess &&
           hipMemcpyAsync(g_buf.h_mm_count, g_buf.d_mismatch_count, sizeof(unsigned int),
                          hipMemcpyDeviceToHost, g_buf.aux) == hipSuccess;
}

bool finish_fixes(const SearchSpec& spec, MismatchIndex& index, Ranker& rk, FixList& fixes) {
    if (hipStreamSynchronize(g_buf.aux) != hipSuccess) return false;
    const unsigned int count = *g_buf.h_mm_count;
    if (count > kMismatchCap) return false;
    std::vector<MismatchChunk>& mm = index.mm;
    mm.resize(count);
    if (count && (hipMemcpyAsync(mm.data(), g_buf.d_mismatch, sizeof(MismatchChunk) * count,
                                 hipMemcpyDeviceToHost, g_buf.aux) != hipSuccess ||
                  hipStreamSynchronize(g_buf.aux) != hipSuccess))
        return false;
    index.sort();
    fixes.clear();
    fixes.reset_failed();
    collect_fixes(spec, index, fixes);
    if (fixes.failed()) return false;
    rk.recount = &index;
    if (fixes.empty()) return true;
    if (!ensure_fixes(fixes.size()) ||
        hipMemcpyAsync(g_buf.d_fixes, fixes.data(), sizeof(CenterFix) * fixes.size(),
                       hipMemcpyHostToDevice, g_buf.aux) != hipSuccess ||
        launch_recount(g_buf.d_fixes, (int32_t)fixes.size(), g_buf.aux) != hipSuccess ||
        hipMemcpyAsync(fixes.data(), g_buf.d_fixes, sizeof(CenterFix) * fixes.size(),
                       hipMemcpyDeviceToHost, g_buf.aux) != hipSuccess ||
        hipStreamSynchronize(g_buf.aux) != hipSuccess)
        return false;
    for (const CenterFix& f : fixes)
        if (f.first_draw >= spec.min_size && f.first_draw <= spec.max_size) rk.pending[f.first_draw]++;
    return true;
}

void wait_while_paused() {
    Control& c = control();
    while (c.pause.load(std::memory_order_relaxed) && !c.cancel.load(std::memory_order_relaxed))
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
}

}  // namespace

void set_max_bands(int nb) { g_max_bands = std::max(1, std::min(kMaxBands, nb)); }

void kernel_stats(int kind, int64_t* launches, int64_t* kernel_ns, int64_t* centers, bool reset) {
    int64_t v[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        // kind 0 plain, 1 banded, 2 both, 3 banded with three x-phases
        if (!(kind == k && k < 2) && !(kind == 2 && k < 2) && !(kind == 3 && k == 2)) continue;
        for (int i = 0; i < 3; ++i) v[i] += g_kstat[k][i].load();
    }
    if (launches) *launches = v[0];
    if (kernel_ns) *kernel_ns = v[1];
    if (centers) *centers = v[2];
    if (reset)
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) g_kstat[k][i].store(0);
}

std::atomic<int> g_last_kernel{-1};

int last_scan_kernel() { return g_last_kernel.load(std::memory_order_relaxed); }

void release_engine() {
    std::lock_guard<std::mutex> lock(g_mutex);
    free_buffers();
}

int64_t run_search(const SearchSpec& spec, ResultRow* out, int32_t max_results,
                   int64_t* hist_out) {
    std::lock_guard<std::mutex> lock(g_mutex);
    Control& ctl = control();
    auto fail = [&]() -> int64_t {
        (void)hipGetLastError();
        ctl.progress.store(0, std::memory_order_relaxed);
        return -1;
    };
    if (!within_search_limit(spec.w, spec.h) || !ensure_buffers()) return fail();

    Ranker rk;
    rk.min_size = spec.min_size;
    rk.max_size = spec.max_size;
    rk.keep = (size_t)max_results;
    rk.emit_min = spec.min_size;

    const bool banded = spec.mode != ScanMode::kPlain;
    // Both kernels count first draws; the centers near a Java-rejected chunk
    // are recounted exactly afterwards.
    std::deque<Work> works;
    FixList& fixes = g_fixes;
    MismatchIndex mismatches;
    if (!start_fixes(spec)) return fail();
    bool fixes_done = false;
    const std::function<bool()> finish = [&]() {
        fixes_done = true;
        return finish_fixes(spec, mismatches, rk, fixes);
    };
    if (banded)
        plan_banded(spec, works);
    else
        plan_plain_rows(spec, spec.x0, spec.w, works);

    int64_t total_centers = 0, done_centers = 0;
    int64_t by_kernel[3] = {0, 0, 0};  // plain, banded one phase, banded two phases
    for (const Work& w : works) {
        total_centers += w.centers;
        by_kernel[w.banded ? (w.ph == 1 ? 1 : 2) : 0] += w.centers;
    }
    if (!works.empty())
        g_last_kernel.store(16 + (int)(std::max_element(by_kernel, by_kernel + 3) - by_kernel),
                            std::memory_order_relaxed);
    bool cancelled = false;
    std::vector<Work> finished;  // launches whose every center is decided
    while (!works.empty()) {
        wait_while_paused();
        if (ctl.cancel.load(std::memory_order_relaxed)) {
            cancelled = true;
            break;
        }
        Work wk = works.front();
        works.pop_front();
        const auto t0 = std::chrono::steady_clock::now();
        // the fixes are finished while the first launch runs
        const RunResult r = fixes_done ? run_work(wk, rk) : run_work(wk, rk, finish);
        if (r == RunResult::kError) return fail();
        if (r == RunResult::kSplit) {
            Work a, b;
            if (!split_work(wk, a, b)) return fail();
            works.push_front(b);
            works.push_front(a);
            continue;
        }
        const auto t1 = std::chrono::steady_clock::now();
        finished.push_back(wk);
        done_centers += wk.centers;
        ctl.processed_centers.fetch_add(wk.centers, std::memory_order_relaxed);
        ctl.scan_work_ns.fetch_add(
            std::max<int64_t>(1, std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count()),
            std::memory_order_relaxed);
        ctl.progress.store(1 + (int32_t)((done_centers * 79) / std::max<int64_t>(1, total_centers)),
                           std::memory_order_relaxed);
    }
    if (!fixes_done && !finish()) return fail();
    // Replace the provisional first-draw counts of recounted centers: all of
    // them, or after a cancel those the finished launches decided (the others
    // were never counted).
    auto decided = [&](const CenterFix& f) {
        if (!cancelled) return true;
        for (const Work& w : finished)
            if (w.banded ? band_tile_of(w.band, w.nb, w.ph, f.cx, f.cz) >= 0
                         : plain_covers(w.plain, f.cx, f.cz))
                return true;
        return false;
    };
    for (const CenterFix& f : fixes) {
        if (!decided(f)) continue;
        if (f.first_draw >= spec.min_size && f.first_draw <= spec.max_size) rk.hist[f.first_draw]--;
        if (f.exact >= spec.min_size && f.exact <= spec.max_size) {
            rk.hist[f.exact]++;
            rk.cand.push_back(Hit{f.exact, f.cx, f.cz, 0});
        }
    }

    int64_t total = 0;
    for (int s = 0; s < kHistBins; ++s) total += rk.hist[s];
    if (hist_out) std::memcpy(hist_out, rk.hist, sizeof(rk.hist));
    write_rows(rk.cand, rk.keep, out);
    ctl.progress.store(90, std::memory_order_relaxed);
    return total;
}

}  // namespace slime
