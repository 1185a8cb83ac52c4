// Multi-search sweep on one device: batched searches (all submitted at once,
// one persistent launch per kernel class) and the served session (continuous
// batching in rounds) share this state machine.
//
// Device-resident per request: its jobs (BandParams pointing at its own
// histogram, emit counter and hit region), per-job claim/done cursors and a
// status word that the persistent kernels read before claiming a tile. Pause,
// resume and cancel are writes of that word; admission reserves a request
// slot, a job range and a hit region whose bytes count against the session's
// HBM budget.
//
// The autotuner: at first use every kernel class is timed on synthetic jobs
// at two tile heights, giving a per-class cost per tile row and a fixed cost
// per tile. Each search's schedule (one or two x-phases, the grouping of its
// z-bands into kernel classes) is the cheapest under that measured model; the
// tile height is the largest that fits the region (at most 2048 slot rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"
#include "search_util.h"
#include "sweep_engine.h"

namespace slime {

namespace {

using Clock = std::chrono::steady_clock;

int64_t ns_since(Clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
}

// P(Bin(221, 0.1) >= t): the chance that one center reaches t.
double center_tail(int t) {
    if (t <= 0) return 1.0;
    double s = 0.0;
    for (int k = t; k <= kDiskCells; ++k)
        s += std::exp(std::lgamma(kDiskCells + 1.0) - std::lgamma(k + 1.0) -
                      std::lgamma(kDiskCells - k + 1.0) + k * std::log(0.1) +
                      (kDiskCells - k) * std::log(0.9));
    return s;
}

constexpr int64_t kMaxHitCap = int64_t(1) << 24;
constexpr int32_t kMaxTileH = 2048;
constexpr int32_t kMinTileH = 16;
constexpr int32_t kTileHalo = 16;  // generated rows beyond the tile's slot rows
// Hits a finishing request may bring back after the device selection (4 MiB).
constexpr uint32_t kSelectCap = 1u << 18;

}  // namespace

// ---- the sweeper -------------------------------------------------------------

// First-fit allocation of index ranges (job slots, hit-region entries).
int64_t Sweeper::Ranges::alloc(int64_t n) {
    if (n <= 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
    }
    return -1;
}

void Sweeper::Ranges::release(int64_t off, int64_t n) {
    if (n <= 0 || off < 0) return;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_.erase(next);
    }
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_.erase(it);
        }
    }
}

Sweeper::~Sweeper() { release(); }

void Sweeper::release() {
    if (device_ >= 0) (void)hipSetDevice(device_);
    // rounds still in flight end first
    while (pend_n_ > 0) {
        (void)hipEventSynchronize(ev_copied_[pend_first_]);
        pend_first_ = (pend_first_ + 1) % kRing;
        --pend_n_;
    }
    pend_first_ = 0;
    last_slot_ = -1;
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    auto ev = [](hipEvent_t& e) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    };
    f(d_jobs_);
    f(d_cursor_);
    f(d_done_);
    f(d_status_);
    f(d_hist_);
    f(d_emitted_);
    f(d_hits_);
    f(d_claims_);
    f(d_table_);
    f(d_queries_);
    f(d_mm_);
    f(d_mm_count_);
    f(d_fixes_);
    f(d_select_);
    f(d_select_count_);
    for (int k = 0; k < kRing; ++k) {
        f(d_order_[k]);
        f(d_first_[k]);
        f(d_start_[k]);
        if (h_done_[k]) (void)hipHostFree(h_done_[k]);
        d_order_[k] = nullptr;
        d_first_[k] = nullptr;
        d_start_[k] = nullptr;
        h_done_[k] = nullptr;
        ev(rev0_[k]);
        ev(rev1_[k]);
        ev(ev_copied_[k]);
        ev(ev_fork_[k]);
        for (int i = 0; i < kLanes - 1; ++i) ev(ev_join_[k][i]);
    }
    ev(ev0_);
    ev(ev1_);
    for (int i = 0; i < kLanes; ++i) {
        if (lane_[i]) (void)hipStreamDestroy(lane_[i]);
        lane_[i] = nullptr;
    }
    if (stream_) (void)hipStreamDestroy(stream_);
    d_jobs_ = nullptr;
    d_cursor_ = d_done_ = d_claims_ = nullptr;
    d_status_ = nullptr;
    d_hist_ = d_emitted_ = nullptr;
    d_hits_ = nullptr;
    d_table_ = nullptr;
    d_queries_ = nullptr;
    d_mm_ = nullptr;
    d_mm_count_ = nullptr;
    d_fixes_ = nullptr;
    d_select_ = nullptr;
    d_select_count_ = nullptr;
    stream_ = nullptr;
    device_ = -1;
    req_cap_ = 0;
    job_cap_ = hit_cap_ = fix_cap_ = 0;
}

bool Sweeper::init(int32_t req_cap, int64_t job_cap, int64_t hit_cap) {
    release();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    device_ = dev;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    cus_ = cus;
    const std::vector<uint64_t>& table = mismatch_table();
    table_n_ = (int32_t)table.size();
    req_cap_ = req_cap;
    job_cap_ = job_cap;
    hit_cap_ = hit_cap;
    fix_cap_ = 1 << 16;
    bool ok =
        hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreate(&ev0_) == hipSuccess && hipEventCreate(&ev1_) == hipSuccess &&
        hipMalloc(&d_jobs_, sizeof(SweepJob) * job_cap) == hipSuccess &&
        hipMalloc(&d_cursor_, sizeof(uint32_t) * job_cap) == hipSuccess &&
        hipMalloc(&d_done_, sizeof(uint32_t) * job_cap) == hipSuccess &&

        hipMalloc(&d_status_, sizeof(int32_t) * req_cap) == hipSuccess &&
        hipMalloc(&d_hist_, sizeof(unsigned long long) * kHistBins * req_cap) == hipSuccess &&
        hipMalloc(&d_emitted_, sizeof(unsigned long long) * req_cap) == hipSuccess &&
        hipMalloc(&d_hits_, sizeof(Hit) * hit_cap) == hipSuccess &&
        hipMalloc(&d_claims_, sizeof(uint32_t) * kSweepKernels) == hipSuccess &&
        hipMalloc(&d_table_, sizeof(uint64_t) * table.size()) == hipSuccess &&
        hipMemcpy(d_table_, table.data(), sizeof(uint64_t) * table.size(),
                  hipMemcpyHostToDevice) == hipSuccess &&
        hipMalloc(&d_queries_, sizeof(MismatchQuery) * req_cap) == hipSuccess &&
        hipMalloc(&d_mm_, sizeof(TaggedMismatch) * kMismatchCap) == hipSuccess &&
        hipMalloc(&d_mm_count_, sizeof(unsigned int)) == hipSuccess &&
        hipMalloc(&d_fixes_, sizeof(CenterFix) * fix_cap_) == hipSuccess &&
        hipMalloc(&d_select_, sizeof(Hit) * kSelectCap) == hipSuccess &&
        hipMalloc(&d_select_count_, sizeof(unsigned int)) == hipSuccess &&
        hipMemsetAsync(d_status_, 0, sizeof(int32_t) * req_cap, stream_) == hipSuccess &&
        hipStreamSynchronize(stream_) == hipSuccess;
    for (int i = 0; ok && i < kLanes; ++i)
        ok = hipStreamCreateWithFlags(&lane_[i], hipStreamNonBlocking) == hipSuccess;
    for (int k = 0; ok && k < kRing; ++k) {
        ok = hipMalloc(&d_order_[k], sizeof(int32_t) * job_cap) == hipSuccess &&
             hipMalloc(&d_first_[k], sizeof(int64_t) * (job_cap + kSweepKernels)) == hipSuccess &&
             hipMalloc(&d_start_[k], sizeof(int32_t) * job_cap) == hipSuccess &&
             hipHostMalloc(&h_done_[k], sizeof(uint32_t) * job_cap, hipHostMallocDefault) ==
                 hipSuccess &&
             hipEventCreate(&rev0_[k]) == hipSuccess && hipEventCreate(&rev1_[k]) == hipSuccess &&
             hipEventCreateWithFlags(&ev_copied_[k], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&ev_fork_[k], hipEventDisableTiming) == hipSuccess;
        for (int i = 0; ok && i < kLanes - 1; ++i)
            ok = hipEventCreateWithFlags(&ev_join_[k][i], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        release();
        return false;
    }
    jobs_alloc_.reset(job_cap);
    hits_alloc_.reset(hit_cap);
    free_slots_.clear();
    for (int32_t s = req_cap - 1; s >= 0; --s) free_slots_.push_back(s);
    for (int k = 0; k < kSweepKernels; ++k) {
        grid_[k] = cus_ * sweep_blocks_per_cu(k);
        direct_grid_[k] = cus_ * batch_blocks_per_cu(k);
    }
    return true;
}

int64_t Sweeper::footprint(const SweepRequest& r) {
    return r.hit_cap * (int64_t)sizeof(Hit) +
           (int64_t)r.jobs.size() * (int64_t)(sizeof(SweepJob) + 3 * sizeof(uint32_t)) +
           (int64_t)(kHistBins + 2) * 8 + (int64_t)r.fixes.size() * (int64_t)sizeof(CenterFix);
}

// ---- autotuner ----

namespace {
// The measured table of each device, shared by every sweeper of the process
// (batches and sessions plan with the same costs, and a new session does not
// measure again).
std::mutex g_cost_mutex;
std::map<int, CostTable> g_cost_tables;
}  // namespace

bool Sweeper::calibrate() {
    if (costs_.ready) return true;
    std::lock_guard<std::mutex> lock(g_cost_mutex);
    const auto known = g_cost_tables.find(device_);
    if (known != g_cost_tables.end()) {
        costs_ = known->second;
        return true;
    }
    // One synthetic job per class and tile height, timed in the one-tile
    // launches that batches and direct rounds run; min_size above any count
    // so nothing is emitted. Slot 0 holds the outputs; its status is set
    // active.
    const int32_t slot = 0;
    const int32_t active = kReqActive;
    if (hipMemcpyAsync(d_status_ + slot, &active, sizeof(int32_t), hipMemcpyHostToDevice,
                       stream_) != hipSuccess)
        return false;
    for (int c = 0; c < kSweepClasses; ++c) {
        const int nb = kSweepClass[c].nb, ph = kSweepClass[c].ph;
        // the kernel searches of the default mode run: the class's union-bound
        // form where it has one
        const int kern = sweep_kernel_of(c, true);
        double a[2] = {0, 0}, r[2] = {0, 0};
        const int32_t ths[2] = {64, 1024};
        for (int k = -1; k < 2; ++k) {  // k = -1: warm-up launch
            const int32_t th = ths[k < 0 ? 0 : k];
            SearchSpec spec;
            spec.seed = 0x5EED5EED5EEDll + c;
            spec.x0 = -kPeriod;
            spec.z0 = 12345;
            spec.w = ph * kPeriod;
            const int64_t tiles_x = kTilesPerPhase;
            const int64_t want = (k < 0 ? 1 : (th == 64 ? 24 : 6)) * (int64_t)direct_grid_[kern];
            const int64_t row_tiles = std::max<int64_t>(1, (want + tiles_x - 1) / tiles_x);
            const int64_t slot_rows = std::min<int64_t>(kPeriod / th * th, row_tiles * th);
            spec.h = (int64_t)(nb - 1) * kPeriod + slot_rows;
            spec.band_rows = (int32_t)slot_rows;
            spec.min_size = kDiskCells;  // count only, emit nothing
            spec.max_size = kDiskCells;
            std::vector<BandJob> jobs;
            band_jobs(spec, x_parts(spec, ph), std::vector<int>{nb}, th, int64_t(1) << 62, jobs);
            if (jobs.size() != 1) return false;
            SweepJob sj{};
            sj.p = jobs[0].p;
            sj.cls = kern;
            sj.req = slot;
            sj.p.tag = slot;
            sj.p.hits = d_hits_;
            sj.p.hit_cap = 0;
            sj.p.emitted = d_emitted_ + slot;
            sj.p.hist = d_hist_ + (size_t)slot * kHistBins;
            const int32_t order0 = 0, start0 = 0;
            const int64_t first[2] = {0, (int64_t)sj.p.tiles_total};
            float ms = 0.f;
            if (hipMemcpyAsync(d_jobs_, &sj, sizeof(sj), hipMemcpyHostToDevice, stream_) !=
                    hipSuccess ||
                hipMemsetAsync(d_done_, 0, sizeof(uint32_t), stream_) != hipSuccess ||
                hipMemcpyAsync(d_order_[0], &order0, sizeof(int32_t), hipMemcpyHostToDevice,
                               stream_) != hipSuccess ||
                hipMemcpyAsync(d_start_[0], &start0, sizeof(int32_t), hipMemcpyHostToDevice,
                               stream_) != hipSuccess ||
                hipMemcpyAsync(d_first_[0], first, sizeof(first), hipMemcpyHostToDevice, stream_) !=
                    hipSuccess ||
                hipEventRecord(ev0_, stream_) != hipSuccess)
                return false;
            const BatchArgs args{d_jobs_, d_order_[0], d_first_[0], d_start_[0], 1, d_done_};
            if (launch_batch(args, kern, first[1], stream_) != hipSuccess ||
                hipEventRecord(ev1_, stream_) != hipSuccess ||
                hipStreamSynchronize(stream_) != hipSuccess ||
                hipEventElapsedTime(&ms, ev0_, ev1_) != hipSuccess)
                return false;
            if (k >= 0) {
                a[k] = (double)ms * 1e6 / (double)sj.p.tiles_total;  // ns per tile
                r[k] = (double)(th + kTileHalo);
            }
        }
        // ns per tile = T * (rows + o)
        double T = (a[1] - a[0]) / (r[1] - r[0]);
        if (!(T > 0)) T = a[1] / r[1];
        const double o = std::max(0.0, a[0] / T - r[0]);
        costs_.ns_row[c] = T;
        costs_.extra_rows[c] = o;
    }
    int32_t zero = 0;
    if (hipMemcpyAsync(d_status_ + slot, &zero, sizeof(int32_t), hipMemcpyHostToDevice, stream_) !=
            hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
        return false;
    costs_.ready = true;
    g_cost_tables[device_] = costs_;
    return true;
}

// Costs are measured per class (with its union-bound kernel where it has
// one); a kernel is costed as its class.
double Sweeper::class_tile_ns(int cls, int32_t tile_h) const {
    return costs_.ns_row[cls] * ((double)tile_h + kTileHalo + costs_.extra_rows[cls]);
}

double Sweeper::tile_ns(int kernel, int32_t tile_h) const {
    return class_tile_ns(sweep_class_of_kernel(kernel), tile_h);
}

// The schedule of one search under the measured costs. Choices:
//  * x: one phase over the whole width; two phases (whole 2P blocks plus a
//    right-aligned block for the remainder, or one masked block between P and
//    2P); two phases where both phases of a slot column are inside and one
//    phase elsewhere (x_parts_one_phase_tail); or, with the union bound and
//    at most 5 bands, three phases past the last whole 2P block;
//  * z, per x-part: the bands grouped into kernel classes (cheapest cover),
//    either all over P slot rows, or with the last full bands split by slot
//    rows around the partial band (best_z below).
bool Sweeper::plan(SweepRequest& r) const {
    const SearchSpec& spec = r.spec;
    struct Group {
        int64_t b_first;
        int nb;
        int32_t s_lo, s_hi;  // slot rows [s_lo, s_hi)
    };
    // the largest tile that fits the rows, but short enough that a small
    // region still has about one tile per CU (a tile's rows are generated one
    // step after another, so few tall tiles bound a small search's time)
    auto tile_for = [&](const XPart& xp, int32_t s_rows) {
        const int64_t col_tiles = (int64_t)xp.nblocks * xp.tiles_per_block();
        const int64_t fit = ((int64_t)s_rows * col_tiles + cus_ - 1) / cus_;
        return (int32_t)std::max<int64_t>(
            kMinTileH, std::min<int64_t>(std::min<int64_t>(kMaxTileH, (s_rows + 3) / 4 * 4),
                                         (fit + 3) / 4 * 4));
    };
    auto group_cost = [&](const XPart& xp, int nb, int32_t s_rows) {
        const int c = sweep_class_of(nb, xp.ph);
        if (c < 0) return -1.0;
        const int32_t th = tile_for(xp, s_rows);
        const int64_t tiles = (int64_t)xp.nblocks * xp.tiles_per_block() * ((s_rows + th - 1) / th);
        return (double)tiles * class_tile_ns(c, th);
    };
    // the union bound has kernels for the classes of up to
    // kUnionBoundMaxBands bands and the three-phase classes: where it
    // applies, only those cover the bands
    const bool ub = union_bound_for(spec.mode, spec.min_size);
    // cheapest cover of bands [b0, b0 + n) over slot rows [s_lo, s_hi)
    auto cover = [&](const XPart& xp, int64_t b0, int64_t n, int32_t s_lo, int32_t s_hi,
                     std::vector<Group>& out) {
        const int32_t s_rows = s_hi - s_lo;
        std::vector<double> cost(n + 1, -1.0);
        std::vector<int> pick(n + 1, 0);
        cost[0] = 0;
        for (int64_t k = 1; k <= n; ++k)
            for (int c = 0; c < kSweepClasses; ++c) {
                const int nb = kSweepClass[c].nb;
                if (!sweep_class_has(c, ub) || kSweepClass[c].ph != xp.ph || nb > k ||
                    cost[k - nb] < 0)
                    continue;
                const double v = cost[k - nb] + group_cost(xp, nb, s_rows);
                if (cost[k] < 0 || v < cost[k]) {
                    cost[k] = v;
                    pick[k] = nb;
                }
            }
        if (cost[n] < 0) return -1.0;
        std::vector<int> sizes;
        for (int64_t k = n; k > 0; k -= pick[k]) sizes.push_back(pick[k]);
        std::sort(sizes.begin(), sizes.end(), std::greater<int>());
        int64_t b = b0;
        for (int nb : sizes) {
            out.push_back(Group{b, nb, s_lo, s_hi});
            b += nb;
        }
        return cost[n];
    };
    // (a) every band over all slot rows; or, with a last, partial band of rem
    // rows after nf full ones, for m = 0..nf: the first nf - m bands over all
    // P rows, the last m full bands and the partial one over rows [0, rem),
    // those m bands alone over rows [rem, P) (m = 0: the partial band as a
    // short group of its own). No tile then scans the partial band's missing
    // rows, and the groups keep high band counts.
    auto best_z = [&](const XPart& xp, std::vector<Group>& out) {
        const int32_t slot_rows = slot_rows_of(spec);
        const int64_t nbz = bands_of(spec);
        std::vector<Group> a;
        double best = cover(xp, 0, nbz, 0, slot_rows, a);
        out = a;
        const int64_t nf = spec.h / kPeriod, rem = spec.h % kPeriod;
        if (spec.band_rows != 0 || nf < 1 || rem <= 0) return best;
        for (int64_t m = 0; m <= nf; ++m) {
            std::vector<Group> g;
            const double c0 = cover(xp, 0, nf - m, 0, kPeriod, g);
            const double c1 = cover(xp, nf - m, m + 1, 0, (int32_t)rem, g);
            const double c2 = cover(xp, nf - m, m, (int32_t)rem, kPeriod, g);
            if (c0 < 0 || c1 < 0 || c2 < 0) continue;
            if (best < 0 || c0 + c1 + c2 < best) {
                best = c0 + c1 + c2;
                out = g;
            }
        }
        return best;
    };
    std::vector<std::vector<XPart>> xopts;
    xopts.push_back(x_parts(spec, 1));
    if (spec.w > kPeriod) xopts.push_back(x_parts(spec, 2));
    if (spec.w > kPeriod && spec.w % (2 * kPeriod) != 0) {
        xopts.push_back(x_parts_one_phase_tail(spec));
        // three phases past the last whole 2P block, for regions of at most 5
        // bands (as the per-search planner: with more, the two-phase parts'
        // wide groups measured faster than the model's three-phase part)
        if (ub && spec.w > 2 * kPeriod && bands_of(spec) <= kThreePhaseMaxBands)
            xopts.push_back(x_parts(spec, 3));
    }
    double best = -1;
    std::vector<XPart> best_parts;
    std::vector<std::vector<Group>> best_groups;
    for (const std::vector<XPart>& parts : xopts) {
        double total = 0;
        std::vector<std::vector<Group>> groups(parts.size());
        for (size_t i = 0; i < parts.size() && total >= 0; ++i) {
            const double c = best_z(parts[i], groups[i]);
            total = c < 0 ? -1 : total + c;
        }
        if (total >= 0 && (best < 0 || total < best)) {
            best = total;
            best_parts = parts;
            best_groups = groups;
        }
    }
    if (best < 0) return false;
    r.jobs.clear();
    for (size_t i = 0; i < best_parts.size(); ++i)
        for (const Group& g : best_groups[i])
            band_jobs_part(spec, best_parts[i], g.b_first, g.nb, g.s_hi,
                           tile_for(best_parts[i], g.s_hi - g.s_lo), int64_t(1) << 62, r.jobs,
                           g.s_lo);
    r.job_cls.clear();
    r.centers = 0;
    for (const BandJob& j : r.jobs) {
        const int c = sweep_class_of(j.nb, j.ph);
        if (c < 0) return false;
        r.job_cls.push_back(sweep_kernel_of(c, ub));
        r.centers += j.centers;
    }
    r.est_ns = best;
    // every hit >= min_size is stored: room for four times the expected count
    const double expect = (double)r.centers * center_tail(spec.min_size);
    r.hit_cap = (int64_t)std::min<double>((double)kMaxHitCap, 4.0 * expect + 8192.0);
    return true;
}

bool Sweeper::reserve(SweepRequest& r) {
    if (free_slots_.empty()) return false;
    const int64_t jo = jobs_alloc_.alloc((int64_t)r.jobs.size());
    if (jo < 0) return false;
    const int64_t ho = hits_alloc_.alloc(r.hit_cap);
    if (ho < 0) {
        jobs_alloc_.release(jo, (int64_t)r.jobs.size());
        return false;
    }
    r.slot = free_slots_.back();
    free_slots_.pop_back();
    r.job_off = jo;
    r.hit_off = ho;
    r.done.assign(r.jobs.size(), 0);
    r.planned.assign(r.jobs.size(), 0);
    r.inflight = 0;
    r.complete = false;
    return true;
}

void Sweeper::unreserve(SweepRequest& r) {
    if (r.slot < 0) return;
    jobs_alloc_.release(r.job_off, (int64_t)r.jobs.size());
    hits_alloc_.release(r.hit_off, r.hit_cap);
    free_slots_.push_back(r.slot);
    r.slot = -1;
    r.job_off = r.hit_off = -1;
}

bool Sweeper::ensure_fix_cap(size_t n) {
    if ((int64_t)n <= fix_cap_) return true;
    if (d_fixes_) (void)hipFree(d_fixes_);
    d_fixes_ = nullptr;
    const int64_t want = std::max<int64_t>((int64_t)n, 2 * fix_cap_);
    if (hipMalloc(&d_fixes_, sizeof(CenterFix) * want) != hipSuccess) {
        fix_cap_ = 0;
        return false;
    }
    fix_cap_ = want;
    return true;
}

bool Sweeper::prepare(const std::vector<SweepRequest*>& reqs) {
    try {
        return prepare_fixes(reqs);
    } catch (const std::bad_alloc&) {
        for (SweepRequest* r : reqs) {
            r->mm.mm.clear();
            r->fixes.clear();
        }
        return false;
    }
}

bool Sweeper::prepare_fixes(const std::vector<SweepRequest*>& reqs) {
    if (reqs.empty()) return true;
    // 1. first-draw mismatches of every region in one launch
    std::vector<MismatchQuery> q;
    int64_t rows = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const SearchSpec& s = reqs[i]->spec;
        MismatchQuery m;
        m.seed = s.seed;
        m.x_first = (int32_t)((int64_t)s.x0 - 8);
        m.z_first = (int32_t)((int64_t)s.z0 - 8);
        m.cols = s.w + 16;
        m.rows = s.h + 16;
        m.row_begin = rows;
        m.tag = (int32_t)i;
        rows += m.rows;
        q.push_back(m);
        reqs[i]->mm.mm.clear();
    }
    for (size_t base = 0; base < q.size(); base += (size_t)req_cap_) {
        const size_t nq = std::min(q.size() - base, (size_t)req_cap_);
        const int64_t first_row = q[base].row_begin;
        std::vector<MismatchQuery> chunk(q.begin() + base, q.begin() + base + nq);
        for (MismatchQuery& m : chunk) m.row_begin -= first_row;
        const int64_t nrows = chunk.back().row_begin + chunk.back().rows;
        unsigned int count = 0;
        if (hipMemcpyAsync(d_queries_, chunk.data(), sizeof(MismatchQuery) * nq,
                           hipMemcpyHostToDevice, stream_) != hipSuccess ||
            hipMemsetAsync(d_mm_count_, 0, sizeof(unsigned int), stream_) != hipSuccess ||
            launch_find_mismatches_multi(d_queries_, (int32_t)nq, nrows, d_table_, table_n_, d_mm_,
                                         kMismatchCap, d_mm_count_, stream_) != hipSuccess ||
            hipMemcpyAsync(&count, d_mm_count_, sizeof(count), hipMemcpyDeviceToHost, stream_) !=
                hipSuccess ||
            hipStreamSynchronize(stream_) != hipSuccess || count > kMismatchCap)
            return false;
        std::vector<TaggedMismatch> mm(count);
        if (count && hipMemcpy(mm.data(), d_mm_, sizeof(TaggedMismatch) * count,
                               hipMemcpyDeviceToHost) != hipSuccess)
            return false;
        for (const TaggedMismatch& t : mm)
            reqs[base + t.tag]->mm.mm.push_back(MismatchChunk{t.x, t.z});
    }
    // 2. the centers they touch, recounted exactly in one launch
    std::vector<CenterFix> all;
    std::vector<size_t> first(reqs.size() + 1, 0);
    for (size_t i = 0; i < reqs.size(); ++i) {
        reqs[i]->mm.sort();
        reqs[i]->fixes.clear();
        collect_fixes(reqs[i]->spec, reqs[i]->mm, reqs[i]->fixes);
        first[i] = all.size();
        all.insert(all.end(), reqs[i]->fixes.begin(), reqs[i]->fixes.end());
    }
    first[reqs.size()] = all.size();
    if (!all.empty()) {
        if (!ensure_fix_cap(all.size()) ||
            hipMemcpyAsync(d_fixes_, all.data(), sizeof(CenterFix) * all.size(),
                           hipMemcpyHostToDevice, stream_) != hipSuccess ||
            launch_recount(d_fixes_, (int32_t)all.size(), stream_) != hipSuccess ||
            hipMemcpyAsync(all.data(), d_fixes_, sizeof(CenterFix) * all.size(),
                           hipMemcpyDeviceToHost, stream_) != hipSuccess ||
            hipStreamSynchronize(stream_) != hipSuccess)
            return false;
        for (size_t i = 0; i < reqs.size(); ++i)
            std::copy(all.begin() + first[i], all.begin() + first[i + 1], reqs[i]->fixes.begin());
    }
    return true;
}

bool Sweeper::upload(const std::vector<SweepRequest*>& reqs) {
    // Jobs, cleared cursors and outputs and the active status of the new
    // requests, written per run of adjacent job ranges and of adjacent slots
    // (a batch's requests form one run each way): a few copies and fill
    // kernels instead of several calls per request.
    std::vector<SweepRequest*> by_job(reqs), by_slot(reqs);
    std::sort(by_job.begin(), by_job.end(),
              [](const SweepRequest* a, const SweepRequest* b) { return a->job_off < b->job_off; });
    std::sort(by_slot.begin(), by_slot.end(),
              [](const SweepRequest* a, const SweepRequest* b) { return a->slot < b->slot; });
    std::vector<SweepJob> staged;
    int64_t run_off = 0;
    auto flush_jobs = [&]() -> bool {
        const size_t nj = staged.size();
        if (!nj) return true;
        const bool ok =
            hipMemcpyAsync(d_jobs_ + run_off, staged.data(), sizeof(SweepJob) * nj,
                           hipMemcpyHostToDevice, stream_) == hipSuccess &&
            launch_fill_u32(d_cursor_ + run_off, nj, 0u, stream_) == hipSuccess &&
            launch_fill_u32(d_done_ + run_off, nj, 0u, stream_) == hipSuccess &&
            hipStreamSynchronize(stream_) == hipSuccess;  // `staged` is reused
        staged.clear();
        return ok;
    };
    for (SweepRequest* r : by_job) {
        if (r->jobs.empty()) continue;
        if (!staged.empty() && r->job_off != run_off + (int64_t)staged.size() && !flush_jobs())
            return false;
        if (staged.empty()) run_off = r->job_off;
        for (size_t j = 0; j < r->jobs.size(); ++j) {
            SweepJob sj{};
            sj.p = r->jobs[j].p;
            sj.cls = r->job_cls[j];
            sj.req = r->slot;
            sj.p.tag = r->slot;
            sj.p.hits = d_hits_ + r->hit_off;
            sj.p.hit_cap = (uint32_t)r->hit_cap;
            sj.p.emitted = d_emitted_ + r->slot;
            sj.p.hist = d_hist_ + (size_t)r->slot * kHistBins;
            staged.push_back(sj);
        }
    }
    if (!flush_jobs()) return false;
    std::vector<int32_t> status;
    int32_t slot0 = 0;
    auto flush_slots = [&]() -> bool {
        const size_t ns = status.size();
        if (!ns) return true;
        const bool ok =
            launch_fill_u32(reinterpret_cast<uint32_t*>(d_hist_ + (size_t)slot0 * kHistBins),
                            2 * (size_t)kHistBins * ns, 0u, stream_) == hipSuccess &&
            launch_fill_u32(reinterpret_cast<uint32_t*>(d_emitted_ + slot0), 2 * ns, 0u,
                            stream_) == hipSuccess &&
            hipMemcpyAsync(d_status_ + slot0, status.data(), sizeof(int32_t) * ns,
                           hipMemcpyHostToDevice, stream_) == hipSuccess &&
            hipStreamSynchronize(stream_) == hipSuccess;  // `status` is reused
        status.clear();
        return ok;
    };
    for (SweepRequest* r : by_slot) {
        if (!status.empty() && r->slot != slot0 + (int32_t)status.size() && !flush_slots())
            return false;
        if (status.empty()) slot0 = r->slot;
        r->status = kReqActive;
        r->remaining_ns = r->est_ns;
        status.push_back(kReqActive);
    }
    return flush_slots();
}

bool Sweeper::clear_status() {
    return hipMemsetAsync(d_status_, 0, sizeof(int32_t) * (size_t)req_cap_, stream_) ==
               hipSuccess &&
           hipStreamSynchronize(stream_) == hipSuccess;
}

bool Sweeper::set_status(SweepRequest& r, int32_t status) {
    r.status = status;
    return hipMemcpyAsync(d_status_ + r.slot, &r.status, sizeof(int32_t), hipMemcpyHostToDevice,
                          stream_) == hipSuccess &&
           hipStreamSynchronize(stream_) == hipSuccess;
}

bool Sweeper::round(const std::vector<SweepRequest*>& reqs, const RoundPolicy& pol,
                    RoundStats* st) {
    while (pend_n_ > 0)
        if (!wait_round(st)) return false;
    const int launched = launch_round(reqs, pol);
    return launched >= 0 && (launched == 0 || wait_round(st));
}

namespace {
bool all_tiles(const SweepRequest& r, const std::vector<uint32_t>& count) {
    for (size_t j = 0; j < r.jobs.size(); ++j)
        if ((int64_t)count[j] < (int64_t)r.jobs[j].p.tiles_total) return false;
    return true;
}
}  // namespace

int Sweeper::launch_round(const std::vector<SweepRequest*>& reqs, const RoundPolicy& pol) {
    if (pend_n_ >= kRing - 1) return -1;  // at most kRing - 1 rounds in flight
    const bool direct = pol.direct;
    if (!direct && pend_n_ > 0) return -1;  // persistent rounds read the device's counts
    // a request with every tile launched and none in flight is complete
    for (SweepRequest* r : reqs)
        if (r->slot >= 0 && r->inflight == 0 && all_tiles(*r, r->done)) r->complete = true;
    // per-class claim order over the jobs of active requests with tiles not
    // yet launched
    struct Item {
        double key;
        int64_t seq;
        int32_t job;
        int64_t left;
        int32_t tile_h;
        int64_t start;  // tiles of the job launched before this round
        SweepRequest* req;
        int32_t j;
    };
    std::vector<Item> items[kSweepKernels];
    int64_t seq = 0;
    for (SweepRequest* r : reqs) {
        if (r->slot < 0 || r->status != kReqActive || r->complete) continue;
        for (size_t j = 0; j < r->jobs.size(); ++j) {
            const int64_t left = (int64_t)r->jobs[j].p.tiles_total - r->planned[j];
            const int c = r->job_cls[j];
            // largest tiles first (batch), or the request with the least work left
            const double key = pol.shortest_first ? r->remaining_ns : -tile_ns(c, r->jobs[j].p.tile_h);
            if (left > 0)
                items[c].push_back(Item{key, seq, (int32_t)(r->job_off + (int64_t)j), left,
                                        r->jobs[j].p.tile_h, (int64_t)r->planned[j], r,
                                        (int32_t)j});
            ++seq;
        }
    }
    const int slot = (pend_first_ + pend_n_) % kRing;
    Pending& pr = pend_[slot];
    pr.order.clear();
    pr.firsts.clear();
    pr.starts.clear();
    pr.expect.clear();
    pr.reqs.clear();
    pr.est_ns = 0;
    std::vector<int32_t>& order = pr.order;
    // Direct rounds: one workgroup per tile of a range of tiles per job (all
    // its tiles left, or as many as the round's device time allows, most
    // urgent jobs first), as the per-search kernels do (their register
    // budget, no claim loop); pause, resume and cancel act between rounds.
    // Otherwise the persistent kernels claim tiles under a claim budget.
    std::vector<int64_t>& firsts = pr.firsts;  // per launch: n + 1 block offsets
    std::vector<int32_t>& starts = pr.starts;  // per direct item: its first tile this round
    struct Launch {
        int cls;
        int32_t off, n;
        uint32_t budget;
        double min_key;
        int32_t first_off;
        int64_t blocks;
        double est_ns;  // modelled device time
    };
    std::vector<Launch> launches;
    // classes holding the most urgent request go first, and the round's device
    // time is handed out in that order: a class takes what its tiles need, up
    // to what is left of round_ns, and at least one tile per resident slot
    // (so every class moves on every round)
    struct ClassWork {
        int c;
        double min_key, per_tile, need_ns;
        int64_t tiles;
    };
    std::vector<ClassWork> work;
    for (int c = 0; c < kSweepKernels; ++c) {
        std::vector<Item>& v = items[c];
        if (v.empty()) continue;
        std::sort(v.begin(), v.end(), [](const Item& a, const Item& b) {
            return a.key != b.key ? a.key < b.key : a.seq < b.seq;
        });
        int64_t tiles = 0;
        double tile_sum = 0;
        for (const Item& it : v) {
            tiles += it.left;
            tile_sum += tile_ns(c, it.tile_h) * (double)it.left;
        }
        work.push_back(ClassWork{c, v.front().key, tile_sum / (double)std::max<int64_t>(1, tiles),
                                 tile_sum, tiles});
    }
    std::stable_sort(work.begin(), work.end(),
                     [](const ClassWork& a, const ClassWork& b) { return a.min_key < b.min_key; });
    double left_ns = (double)pol.round_ns;
    for (const ClassWork& cw : work) {
        const int c = cw.c;
        std::vector<Item>& v = items[c];
        Launch l;
        l.cls = c;
        l.off = (int32_t)order.size();
        l.min_key = cw.min_key;
        int64_t want = INT64_MAX;
        if (pol.round_ns > 0) {
            // a direct launch takes whole waves of its resident workgroups, so
            // that its tiles end together (a last partial wave would hold the
            // device for a whole tile at part occupancy)
            const double g = (double)(direct ? direct_grid_[c] : grid_[c]);
            const double per_tile = std::max(1e-3, cw.per_tile);
            const double give = std::max(std::min(cw.need_ns, left_ns), per_tile * g);
            double tiles = give / per_tile;
            if (direct) tiles = std::max(1.0, std::round(tiles / g)) * g;
            left_ns -= std::min(tiles * per_tile, cw.need_ns);
            want = (int64_t)std::min<double>(4e9, std::max(g, tiles));
        }
        l.budget = pol.round_ns > 0 ? (uint32_t)want : 0xFFFFFFFFu;
        l.first_off = (int32_t)firsts.size();
        l.blocks = 0;
        const double est_before = pr.est_ns;
        if (direct) {
            // a launch's grid holds at most 2^31 - 1 workgroups; the rest of
            // the tiles wait for the next round
            const int64_t cap = std::min<int64_t>(want, 0x7FFFFFFF);
            firsts.push_back(0);
            for (const Item& it : v) {
                const int64_t n = std::min<int64_t>(it.left, cap - l.blocks);
                if (n <= 0) break;
                order.push_back(it.job);
                starts.push_back((int32_t)it.start);
                l.blocks += n;
                firsts.push_back(l.blocks);
                pr.expect.push_back(Expect{it.req, it.j, (uint32_t)(it.start + n)});
                pr.est_ns += (double)n * tile_ns(c, it.tile_h);
            }
        } else {
            for (const Item& it : v) order.push_back(it.job);
            pr.est_ns += std::min(cw.need_ns, (double)want * cw.per_tile);
        }
        l.n = (int32_t)order.size() - l.off;
        l.est_ns = pr.est_ns - est_before;
        launches.push_back(l);
    }
    if (launches.empty()) return 0;
    // Direct launches of several classes run side by side on the lane
    // streams (they touch disjoint jobs; a request's counters are atomics), so
    // that one launch's last workgroups overlap the next one's; the round
    // ends when all have ended. A captured graph keeps one stream.
    // A round launched while the previous one runs starts on the lane after
    // the one holding the previous round's longest launch, so that its most
    // urgent launch does not queue behind that one.
    const int lanes = direct && !pol.use_graph ? std::min<int>(kLanes, (int)launches.size()) : 1;
    const int rot = direct && !pol.use_graph ? (big_lane_ + 1) % kLanes : 0;
    hipStream_t join = lane_[rot];  // the round's copies, events and the join run there
    if (hipMemcpyAsync(d_order_[slot], order.data(), sizeof(int32_t) * order.size(),
                       hipMemcpyHostToDevice, join) != hipSuccess ||
        (direct && (hipMemcpyAsync(d_first_[slot], firsts.data(), sizeof(int64_t) * firsts.size(),
                                   hipMemcpyHostToDevice, join) != hipSuccess ||
                    hipMemcpyAsync(d_start_[slot], starts.data(), sizeof(int32_t) * starts.size(),
                                   hipMemcpyHostToDevice, join) != hipSuccess)))
        return -1;
    auto lane_of = [&](size_t li) { return lane_[(rot + (int)(li % (size_t)lanes)) % kLanes]; };
    auto enqueue = [&]() -> bool {
        if (launch_fill_u32(d_claims_, kSweepKernels, 0u, join) != hipSuccess) return false;
        if (lanes > 1 && hipEventRecord(ev_fork_[slot], join) != hipSuccess) return false;
        for (int i = 1; i < lanes; ++i)
            if (hipStreamWaitEvent(lane_of((size_t)i), ev_fork_[slot], 0) != hipSuccess)
                return false;
        for (size_t li = 0; li < launches.size(); ++li) {
            const Launch& l = launches[li];
            if (direct) {
                BatchArgs a{d_jobs_, d_order_[slot] + l.off, d_first_[slot] + l.first_off,
                            d_start_[slot] + l.off, l.n, d_done_};
                if (launch_batch(a, l.cls, l.blocks, lane_of(li)) != hipSuccess) return false;
                continue;
            }
            SweepArgs a{d_jobs_, d_order_[slot] + l.off, l.n, d_cursor_, d_done_, d_status_,
                        d_claims_ + l.cls, l.budget};
            if (launch_sweep(a, l.cls, grid_[l.cls], join) != hipSuccess) return false;
        }
        for (int i = 1; i < lanes; ++i)
            if (hipEventRecord(ev_join_[slot][i - 1], lane_of((size_t)i)) != hipSuccess ||
                hipStreamWaitEvent(join, ev_join_[slot][i - 1], 0) != hipSuccess)
                return false;
        return true;
    };
    if (hipEventRecord(rev0_[slot], join) != hipSuccess) return -1;
    if (pol.use_graph) {
        // the launch sequence, keyed by its arguments; replays reuse the device
        // arrays (jobs, orders) whose contents were written above
        std::vector<int64_t> key;
        key.push_back(direct ? 1 : 0);
        key.push_back(slot);
        for (const Launch& l : launches) {
            key.push_back(l.cls);
            key.push_back(l.off);
            key.push_back(l.n);
            key.push_back(l.budget);
            key.push_back(l.first_off);
            key.push_back(l.blocks);
        }
        auto it = graphs_.find(key);
        if (it == graphs_.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            if (hipStreamBeginCapture(join, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return -1;
            const bool ok = enqueue();
            if (hipStreamEndCapture(join, &g) != hipSuccess || !ok) {
                if (g) (void)hipGraphDestroy(g);
                return -1;
            }
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return -1;
            it = graphs_.emplace(key, ge).first;
        }
        if (hipGraphLaunch(it->second, join) != hipSuccess) return -1;
    } else if (!enqueue()) {
        return -1;
    }
    if (hipEventRecord(rev1_[slot], join) != hipSuccess ||
        hipMemcpyAsync(h_done_[slot], d_done_, sizeof(uint32_t) * (size_t)job_cap_,
                       hipMemcpyDeviceToHost, join) != hipSuccess ||
        hipEventRecord(ev_copied_[slot], join) != hipSuccess)
        return -1;
    size_t big = 0;
    for (size_t li = 1; li < launches.size(); ++li)
        if (launches[li].est_ns > launches[big].est_ns) big = li;
    big_lane_ = (rot + (int)(big % (size_t)lanes)) % kLanes;
    pr.direct = direct;
    pr.launches = (int32_t)launches.size();
    for (const Expect& e : pr.expect) {
        SweepRequest& r = *e.req;
        r.planned[e.j] = e.count;
    }
    for (SweepRequest* r : reqs) {
        bool in = !direct && r->slot >= 0 && r->status == kReqActive && !r->complete;
        if (direct)
            for (const Expect& e : pr.expect)
                if (e.req == r) {
                    in = true;
                    break;
                }
        if (!in) continue;
        ++r->inflight;
        pr.reqs.push_back(r);
        double left_ns = 0;
        for (size_t j = 0; j < r->jobs.size(); ++j) {
            const int64_t left = (int64_t)r->jobs[j].p.tiles_total - r->planned[j];
            if (left > 0) left_ns += (double)left * tile_ns(r->job_cls[j], r->jobs[j].p.tile_h);
        }
        r->remaining_ns = left_ns;
    }
    const Clock::time_point now = Clock::now();
    Clock::time_point begin = now;
    if (pend_n_ > 0) begin = std::max(begin, pend_[(pend_first_ + pend_n_ - 1) % kRing].t_end);
    pr.t_launch = now;
    pr.t_end = begin + std::chrono::nanoseconds((int64_t)pr.est_ns);
    ++pend_n_;
    return 1;
}

bool Sweeper::wait_round(RoundStats* st) {
    if (pend_n_ == 0) return true;
    const int slot = pend_first_;
    Pending& pr = pend_[slot];
    if (hipEventSynchronize(ev_copied_[slot]) != hipSuccess) return false;
    pend_first_ = (pend_first_ + 1) % kRing;
    --pend_n_;
    const Clock::time_point now = Clock::now();
    if (pend_n_ > 0) {  // the next round runs from about now
        Pending& nx = pend_[pend_first_];
        nx.t_end = std::max(now, nx.t_launch) + std::chrono::nanoseconds((int64_t)nx.est_ns);
    }
    // device time: the round's span less what overlapped the previous round
    float ms = 0.f, before = 0.f;
    if (hipEventElapsedTime(&ms, rev0_[slot], rev1_[slot]) != hipSuccess) return false;
    if (last_slot_ >= 0 && hipEventElapsedTime(&before, rev0_[slot], rev1_[last_slot_]) == hipSuccess)
        ms -= std::min(ms, std::max(0.f, before));
    else
        (void)hipGetLastError();
    last_slot_ = slot;
    if (st) {
        st->kernel_ns += (int64_t)((double)ms * 1e6);
        st->launches += pr.launches;
        st->rounds += 1;
    }
    const uint32_t* h = h_done_[slot];
    if (pr.direct) {
        // the tiles of this round are done (later rounds may have added more)
        for (const Expect& e : pr.expect) {
            SweepRequest& r = *e.req;
            if (h[r.job_off + e.j] < e.count) return false;
            r.done[e.j] = std::max(r.done[e.j], e.count);
        }
    } else {
        for (SweepRequest* r : pr.reqs)
            for (size_t j = 0; j < r->jobs.size(); ++j) {
                r->done[j] = h[r->job_off + (int64_t)j];
                r->planned[j] = r->done[j];
            }
    }
    for (SweepRequest* r : pr.reqs) {
        --r->inflight;
        if (!pr.direct) {
            double left_ns = 0;
            for (size_t j = 0; j < r->jobs.size(); ++j) {
                const int64_t left = (int64_t)r->jobs[j].p.tiles_total - r->done[j];
                if (left > 0) left_ns += (double)left * tile_ns(r->job_cls[j], r->jobs[j].p.tile_h);
            }
            r->remaining_ns = left_ns;
        }
        if (r->inflight == 0 && all_tiles(*r, r->done)) r->complete = true;
    }
    return true;
}

double Sweeper::progress(const SweepRequest& r) const {
    int64_t done = 0, total = 0;
    for (size_t j = 0; j < r.jobs.size(); ++j) {
        done += r.done[j];
        total += r.jobs[j].p.tiles_total;
    }
    return total ? (double)done / (double)total : 1.0;
}

bool Sweeper::rank_begin(SweepRequest& r, bool partial, const unsigned long long* hist,
                         unsigned long long emitted, RankState& rs) {
    if ((int64_t)emitted > r.hit_cap && !partial) {
        // more hits than the reserved region: rerun alone with a rising gate
        r.rows.assign((size_t)r.keep, ResultRow{});
        r.total = run_search(r.spec, r.rows.data(), r.keep);
        r.rows.resize((size_t)std::max<int64_t>(0, std::min<int64_t>(r.total, r.keep)));
        r.fallback = true;
        rs.done = true;
        return r.total >= 0;
    }
    // A cancelled request keeps what its finished tiles found: the total
    // counts every decided center; the rows rank the stored hits (the first
    // hit_cap emitted if the region overflowed: then r.truncated).
    rs.stored = (size_t)std::min<int64_t>((int64_t)emitted, r.hit_cap);
    r.truncated = (int64_t)emitted > r.hit_cap;
    int64_t h[kHistBins];
    for (int s = 0; s < kHistBins; ++s) h[s] = (int64_t)hist[s];
    // Replace the provisional first-draw counts of recounted centers: all of
    // them, or for a cancelled request those in finished tiles (every job's
    // tiles finish in index order, so its first done[j] tiles).
    auto decided = [&](const CenterFix& f) {
        if (!partial) return true;
        for (size_t j = 0; j < r.jobs.size(); ++j) {
            const int64_t t = band_tile_of(r.jobs[j].p, r.jobs[j].nb, r.jobs[j].ph, f.cx, f.cz);
            if (t >= 0) return t < (int64_t)r.done[j];
        }
        return false;
    };
    for (const CenterFix& f : r.fixes) {
        if (!decided(f)) continue;
        if (f.first_draw >= r.spec.min_size && f.first_draw <= r.spec.max_size) h[f.first_draw]--;
        if (f.exact >= r.spec.min_size && f.exact <= r.spec.max_size) {
            h[f.exact]++;
            rs.cand.push_back(Hit{f.exact, f.cx, f.cz, 0});
        }
    }
    int64_t total = 0;
    for (int s = 0; s < kHistBins; ++s) total += h[s];
    r.total = total;
    // Top-K on the device: with the corrected histogram, the best `keep`
    // centers all score at least t, the highest score with >= keep centers at
    // or above it. Only the stored hits of score >= t come back (those near a
    // Java-rejected chunk are dropped below and replaced by the fixes, which
    // the histogram already counts). A truncated region, or more such hits
    // than the staging buffer holds, brings every stored hit back.
    rs.t = r.spec.min_size;
    for (int64_t s = kHistBins - 1, acc = 0; s > r.spec.min_size; --s) {
        acc += h[s];
        if (acc >= (int64_t)r.keep) {
            rs.t = (int32_t)s;
            break;
        }
    }
    rs.select = !r.truncated && rs.t > r.spec.min_size && rs.stored > (size_t)r.keep;
    return true;
}

bool Sweeper::rank_hits(SweepRequest& r, const RankState& rs, std::vector<Hit>& hits) {
    if (rs.select) {
        unsigned int n_sel = 0;
        if (launch_fill_u32(d_select_count_, 1, 0u, stream_) != hipSuccess ||
            launch_select_hits(d_hits_ + r.hit_off, (int64_t)rs.stored, rs.t, d_select_,
                               kSelectCap, d_select_count_, stream_) != hipSuccess ||
            hipMemcpyAsync(&n_sel, d_select_count_, sizeof(n_sel), hipMemcpyDeviceToHost,
                           stream_) != hipSuccess ||
            hipStreamSynchronize(stream_) != hipSuccess)
            return false;
        if (n_sel <= kSelectCap) {
            hits.resize(n_sel);
            return !n_sel || (hipMemcpyAsync(hits.data(), d_select_, sizeof(Hit) * n_sel,
                                             hipMemcpyDeviceToHost, stream_) == hipSuccess &&
                              hipStreamSynchronize(stream_) == hipSuccess);
        }
    }
    hits.resize(rs.stored);
    return !rs.stored || (hipMemcpyAsync(hits.data(), d_hits_ + r.hit_off, sizeof(Hit) * rs.stored,
                                         hipMemcpyDeviceToHost, stream_) == hipSuccess &&
                          hipStreamSynchronize(stream_) == hipSuccess);
}

void Sweeper::rank_end(SweepRequest& r, RankState& rs, const std::vector<Hit>& hits) {
    r.hits_stored = (int64_t)rs.stored;
    r.hits_copied = (int64_t)hits.size();
    std::vector<Hit>& cand = rs.cand;
    cand.reserve(cand.size() + hits.size());
    for (const Hit& x : hits)
        if (!r.mm.affects(x.cx, x.cz)) cand.push_back(Hit{x.size, x.cx, x.cz, 0});
    r.rows.assign((size_t)std::min<int64_t>((int64_t)cand.size(), r.keep), ResultRow{});
    write_rows(cand, (size_t)r.keep, r.rows.data());
}

bool Sweeper::finalize(SweepRequest& r, bool partial) {
    unsigned long long hist[kHistBins];
    unsigned long long emitted = 0;
    if (hipMemcpyAsync(hist, d_hist_ + (size_t)r.slot * kHistBins, sizeof(hist),
                       hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipMemcpyAsync(&emitted, d_emitted_ + r.slot, sizeof(emitted), hipMemcpyDeviceToHost,
                       stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
        return false;
    RankState rs;
    if (!rank_begin(r, partial, hist, emitted, rs)) return false;
    if (rs.done) return true;
    std::vector<Hit> hits;
    if (!rank_hits(r, rs, hits)) return false;
    rank_end(r, rs, hits);
    return true;
}

bool Sweeper::finalize_all(const std::vector<SweepRequest*>& rs_in, bool partial) {
    if (rs_in.size() <= 1) {
        for (SweepRequest* r : rs_in)
            if (!finalize(*r, partial)) return false;
        return true;
    }
    // every slot's counts in two copies, one synchronisation
    std::vector<unsigned long long> hist((size_t)req_cap_ * kHistBins), emitted((size_t)req_cap_);
    if (hipMemcpyAsync(hist.data(), d_hist_, sizeof(unsigned long long) * hist.size(),
                       hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipMemcpyAsync(emitted.data(), d_emitted_, sizeof(unsigned long long) * emitted.size(),
                       hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
        return false;
    std::vector<RankState> st(rs_in.size());
    std::vector<std::vector<Hit>> hits(rs_in.size());
    bool copies = false;
    for (size_t i = 0; i < rs_in.size(); ++i) {
        SweepRequest& r = *rs_in[i];
        if (!rank_begin(r, partial, hist.data() + (size_t)r.slot * kHistBins, emitted[r.slot],
                        st[i]))
            return false;
        if (st[i].done) continue;
        if (st[i].select) {
            if (!rank_hits(r, st[i], hits[i])) return false;  // selections one at a time
        } else if (st[i].stored) {
            // the stored hits of every other request in one more round trip
            hits[i].resize(st[i].stored);
            if (hipMemcpyAsync(hits[i].data(), d_hits_ + r.hit_off, sizeof(Hit) * st[i].stored,
                               hipMemcpyDeviceToHost, stream_) != hipSuccess)
                return false;
            copies = true;
        }
    }
    if (copies && hipStreamSynchronize(stream_) != hipSuccess) return false;
    for (size_t i = 0; i < rs_in.size(); ++i)
        if (!st[i].done) rank_end(*rs_in[i], st[i], hits[i]);
    return true;
}

// ---- batch ----------------------------------------------------------------------

namespace {
std::mutex g_batch_mutex;
Sweeper g_batch_sweeper;
BatchStats g_last_batch;
}  // namespace

int run_batch(const SearchSpec* specs, int32_t n, int32_t max_results, ResultRow* out,
              int64_t* totals, const SweepOptions& opt, BatchStats* stats) {
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    const auto t0 = Clock::now();
    BatchStats st{};
    Sweeper& sw = g_batch_sweeper;
    std::vector<SweepRequest> reqs((size_t)n);
    int64_t need_jobs = 0, need_hits = 0;
    auto fail = [&]() {
        (void)hipGetLastError();
        return -1;
    };
    // capacity first: plans only need the cost table, which needs a device state
    if (!sw.ready() && !sw.init(std::max(n, 1), 4096, int64_t(1) << 20)) return fail();
    if (!sw.calibrate()) return fail();
    for (int32_t i = 0; i < n; ++i) {
        reqs[i].spec = specs[i];
        reqs[i].keep = max_results;
        if (!sw.plan(reqs[i])) return fail();
        need_jobs += (int64_t)reqs[i].jobs.size();
        need_hits += reqs[i].hit_cap;
    }
    if (sw.req_cap() < n || sw.job_cap() < need_jobs || sw.hit_cap() < need_hits) {
        const CostTable keep = sw.costs();
        if (!sw.init(std::max(n, sw.req_cap()), std::max(need_jobs, sw.job_cap()),
                     std::max(need_hits, sw.hit_cap())))
            return fail();
        sw.set_costs(keep);
    }
    std::vector<SweepRequest*> ptrs;
    for (SweepRequest& r : reqs) {
        if (!sw.reserve(r)) return fail();
        ptrs.push_back(&r);
    }
    auto cleanup = [&]() {
        (void)sw.clear_status();
        for (SweepRequest& r : reqs) sw.unreserve(r);
    };
    if (!sw.prepare(ptrs) || !sw.upload(ptrs)) {
        cleanup();
        return fail();
    }
    st.prepare_ns = ns_since(t0);
    RoundPolicy pol;
    pol.round_ns = 0;  // one unbounded round: every tile in one launch per class
    pol.shortest_first = false;
    pol.direct = !opt.persistent;
    pol.use_graph = opt.use_graph;
    RoundStats rs{};
    const auto t1 = Clock::now();
    bool ok = sw.round(ptrs, pol, &rs);
    for (int guard = 0; ok && guard < 4; ++guard) {  // only after a budget cut-off
        bool all = true;
        for (SweepRequest* r : ptrs) all = all && r->complete;
        if (all) break;
        ok = sw.round(ptrs, pol, &rs);
    }
    st.sweep_ns = ns_since(t1);
    if (!ok) {
        cleanup();
        return fail();
    }
    for (SweepRequest* r : ptrs)
        if (!r->complete) {
            cleanup();
            return fail();
        }
    if (!sw.finalize_all(ptrs, false)) {
        cleanup();
        return fail();
    }
    for (int32_t i = 0; i < n; ++i) {
        SweepRequest& r = reqs[i];
        totals[i] = r.total;
        std::memcpy(out + (size_t)i * max_results, r.rows.data(),
                    sizeof(ResultRow) * r.rows.size());
        st.centers += r.centers;
        st.fallbacks += r.fallback ? 1 : 0;
        st.jobs += (int32_t)r.jobs.size();
        st.hits_stored += r.hits_stored;
        st.hits_copied += r.hits_copied;
        for (size_t j = 0; j < r.jobs.size(); ++j)
            st.class_centers[sweep_class_of_kernel(r.job_cls[j])] += r.jobs[j].centers;
    }
    cleanup();
    st.kernel_ns = rs.kernel_ns;
    st.launches = rs.launches;
    st.wall_ns = ns_since(t0);
    g_last_batch = st;
    if (stats) *stats = st;
    return 0;
}

void last_batch_stats(BatchStats* st) {
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    if (st) *st = g_last_batch;
}

bool sweep_cost_table(double* ns_row, double* extra_rows) {
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    Sweeper& sw = g_batch_sweeper;
    if (!sw.ready() && !sw.init(16, 4096, int64_t(1) << 20)) return false;
    if (!sw.calibrate()) return false;
    for (int c = 0; c < kSweepClasses; ++c) {
        if (ns_row) ns_row[c] = sw.costs().ns_row[c];
        if (extra_rows) extra_rows[c] = sw.costs().extra_rows[c];
    }
    return true;
}

}  // namespace slime

// ---- served session --------------------------------------------------------------

namespace slime {

struct Session::Entry {
    int64_t id = 0;
    SweepRequest req;
    ReqState state = ReqState::kQueued;
    bool want_pause = false, want_resume = false, want_cancel = false;
    Clock::time_point t_submit, t_done;
    int64_t footprint = 0;
    double progress = 0;
};

Session::Session(int64_t hbm_budget, int32_t max_active, int32_t round_ms, bool persistent) {
    persistent_ = persistent;
    budget_ = std::max<int64_t>(hbm_budget, int64_t(64) << 20);
    max_active_ = std::max(1, max_active);
    round_ns_ = (int64_t)std::max(1, round_ms) * 1000000;
    // hit regions take most of the budget; jobs are a few per request
    const int64_t hits = std::max<int64_t>(int64_t(1) << 16, budget_ / (int64_t)sizeof(Hit));
    if (!sw_.init(max_active_, 64 * (int64_t)max_active_ + 4096, hits) || !sw_.calibrate()) return;
    st_.hbm_budget = budget_;
    ok_ = true;
    thread_ = std::thread([this] { worker(); });
}

Session::~Session() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
}

int64_t Session::submit(const SearchSpec& spec, int32_t keep) {
    if (!ok_ || keep <= 0 || !within_search_limit(spec.w, spec.h)) return -1;
    std::unique_ptr<Entry> e;
    try {
        e = std::make_unique<Entry>();
        e->req.spec = spec;
        e->req.keep = keep;
        e->t_submit = Clock::now();
        if (!sw_.plan(e->req)) return -1;  // reads only the cost table
    } catch (const std::bad_alloc&) {
        return -1;
    }
    e->footprint = Sweeper::footprint(e->req);
    // requests that could never be admitted: over the HBM budget, or more
    // jobs or hits than the session's device arrays hold
    if (e->footprint > budget_ || (int64_t)e->req.jobs.size() > sw_.job_cap() ||
        e->req.hit_cap > sw_.hit_cap())
        return -1;
    std::lock_guard<std::mutex> lk(mu_);
    const int64_t id = next_id_++;
    e->id = id;
    entries_[id] = std::move(e);
    queue_.push_back(id);
    cv_.notify_all();
    return id;
}

bool Session::pause(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = entries_.find(id);
    if (it == entries_.end()) return false;
    Entry& e = *it->second;
    if (e.state != ReqState::kQueued && e.state != ReqState::kRunning) return false;
    e.want_pause = true;
    e.want_resume = false;
    cv_.notify_all();
    return true;
}

bool Session::resume(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = entries_.find(id);
    if (it == entries_.end()) return false;
    Entry& e = *it->second;
    if (e.state != ReqState::kPaused && !e.want_pause) return false;
    e.want_resume = true;
    e.want_pause = false;
    cv_.notify_all();
    return true;
}

bool Session::cancel(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = entries_.find(id);
    if (it == entries_.end()) return false;
    Entry& e = *it->second;
    if (e.state == ReqState::kDone || e.state == ReqState::kCancelled ||
        e.state == ReqState::kFailed)
        return false;
    e.want_cancel = true;
    cv_.notify_all();
    return true;
}

bool Session::poll(int64_t id, ReqState* state, double* progress, int64_t* total,
                   int64_t* latency_ns) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = entries_.find(id);
    if (it == entries_.end()) return false;
    const Entry& e = *it->second;
    const bool fin = e.state == ReqState::kDone || e.state == ReqState::kCancelled;
    if (state) *state = e.state;
    if (progress) *progress = e.state == ReqState::kDone ? 1.0 : e.progress;
    if (total) *total = fin ? e.req.total : -1;
    if (latency_ns)
        *latency_ns = fin ? std::chrono::duration_cast<std::chrono::nanoseconds>(e.t_done -
                                                                                 e.t_submit)
                                .count()
                          : -1;
    return true;
}

ReqState Session::wait(int64_t id, int32_t timeout_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto finished = [&] {
        auto it = entries_.find(id);
        if (it == entries_.end()) return true;
        const ReqState s = it->second->state;
        return s == ReqState::kDone || s == ReqState::kCancelled || s == ReqState::kFailed;
    };
    if (timeout_ms < 0)
        done_cv_.wait(lk, finished);
    else
        done_cv_.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished);
    auto it = entries_.find(id);
    return it == entries_.end() ? ReqState::kFailed : it->second->state;
}

int64_t Session::result(int64_t id, ResultRow* out, int32_t cap) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = entries_.find(id);
    if (it == entries_.end()) return -1;
    const Entry& e = *it->second;
    if (e.state != ReqState::kDone && e.state != ReqState::kCancelled) return -1;
    const size_t n = std::min(e.req.rows.size(), (size_t)std::max(0, cap));
    if (out && n) std::memcpy(out, e.req.rows.data(), sizeof(ResultRow) * n);
    return e.req.total;
}

bool Session::release(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = entries_.find(id);
    if (it == entries_.end()) return false;
    const ReqState s = it->second->state;
    if (s != ReqState::kDone && s != ReqState::kCancelled && s != ReqState::kFailed) return false;
    entries_.erase(it);
    return true;
}

SessionStats Session::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    SessionStats s = st_;
    s.active = (int32_t)admitted_.size();
    s.queued = (int32_t)queue_.size();
    s.hbm_in_use = in_use_;
    return s;
}

void Session::worker() {
    std::unique_lock<std::mutex> lk(mu_);
    // Ranks the finished requests (every tile done) and the cancelled ones,
    // then releases them; called with the lock held, drops it for the device
    // work. ok false: a device error fails every admitted request (queued
    // ones wait).
    auto finish = [&](bool ok, const std::vector<Entry*>& cancelled, const RoundStats& rs) {
        auto is_cancelled = [&](const Entry& e) {
            return std::find(cancelled.begin(), cancelled.end(), &e) != cancelled.end();
        };
        std::vector<SweepRequest*> complete;
        for (int64_t id : admitted_) {
            Entry& e = *entries_[id];
            if (e.state == ReqState::kRunning && e.req.complete && !is_cancelled(e))
                complete.push_back(&e.req);
        }
        lk.unlock();
        if (ok) ok = sw_.finalize_all(complete, false);
        for (Entry* e : cancelled) ok = ok && sw_.finalize(e->req, true);
        if (!ok)  // nothing stays in flight over the requests about to be released
            while (sw_.in_flight() > 0 && sw_.wait_round(nullptr)) {
            }
        lk.lock();
        st_.kernel_ns += rs.kernel_ns;
        st_.rounds += rs.rounds;
        const auto now = Clock::now();
        for (auto it = admitted_.begin(); it != admitted_.end();) {
            Entry& e = *entries_[*it];
            e.progress = sw_.progress(e.req);
            const bool was_cancelled = is_cancelled(e);
            const bool done = e.state == ReqState::kRunning && e.req.complete;
            if (!ok || done || was_cancelled) {
                e.state = !ok ? ReqState::kFailed
                              : was_cancelled ? ReqState::kCancelled : ReqState::kDone;
                e.t_done = now;
                if (done && ok) st_.centers_done += e.req.centers;
                sw_.unreserve(e.req);
                in_use_ -= e.footprint;
                admit_blocked_ = false;
                it = admitted_.erase(it);
                continue;
            }
            ++it;
        }
        done_cv_.notify_all();
        if (!ok) (void)hipGetLastError();
    };
    for (;;) {
        auto has_work = [&] {
            if (stop_) return true;
            if (!queue_.empty() && !admit_blocked_ &&
                (int32_t)admitted_.size() < max_active_ &&
                in_use_ + entries_[queue_.front()]->footprint <= budget_)
                return true;
            for (int64_t id : queue_)
                if (entries_[id]->want_cancel) return true;
            for (int64_t id : admitted_) {
                const Entry& e = *entries_[id];
                if (e.state == ReqState::kRunning || e.want_cancel || e.want_resume ||
                    e.want_pause)
                    return true;
            }
            return sw_.in_flight() > 0;
        };
        cv_.wait(lk, has_work);
        if (stop_) break;
        // Direct rounds are pipelined: with one round in flight the worker
        // waits until shortly before that round is modelled to end, then
        // admits, applies controls and launches the next round, which the
        // device starts as the current one drains (the host's work between
        // rounds and the last launch's tail are hidden), and only then waits
        // for the current round. The model's end can be late (a round that
        // ends early would leave the device idle): the worker looks at the
        // round every half millisecond, and a round found ended has its
        // requests finished first (before new requests are prepared).
        if (!persistent_ && sw_.in_flight() == 1) {
            const auto margin =
                std::chrono::nanoseconds(std::clamp<int64_t>(round_ns_ / 8, 1000000, 3000000));
            const auto due = sw_.expected_end() - margin;
            while (!stop_ && Clock::now() < due && !sw_.oldest_ended())
                cv_.wait_until(lk, std::min(due, Clock::now() + std::chrono::microseconds(500)),
                               [&] { return stop_; });
            if (stop_) break;
            if (sw_.oldest_ended()) {
                RoundStats rs0{};
                lk.unlock();
                const bool ok0 = sw_.wait_round(&rs0);
                lk.lock();
                finish(ok0, {}, rs0);
            }
        }

        // 1. control edits of queued requests, then admission in arrival order
        for (auto qi = queue_.begin(); qi != queue_.end();) {
            Entry& e = *entries_[*qi];
            if (e.want_cancel) {
                e.want_cancel = false;
                e.state = ReqState::kCancelled;
                e.req.total = 0;
                e.t_done = Clock::now();
                qi = queue_.erase(qi);
                done_cv_.notify_all();
                continue;
            }
            ++qi;
        }
        std::vector<Entry*> admit;
        while (!queue_.empty() && (int32_t)admitted_.size() < max_active_) {
            Entry& e = *entries_[queue_.front()];
            if (in_use_ + e.footprint > budget_) break;
            if (!sw_.reserve(e.req)) {  // fragmented: wait until a request leaves
                admit_blocked_ = true;
                break;
            }
            in_use_ += e.footprint;
            admitted_.push_back(queue_.front());
            queue_.pop_front();
            admit.push_back(&e);
        }
        // status edits of admitted requests (applied on the device below)
        std::vector<std::pair<Entry*, int32_t>> edits;
        std::vector<Entry*> cancelled;
        for (int64_t id : admitted_) {
            Entry& e = *entries_[id];
            if (e.want_cancel) {
                e.want_cancel = false;
                cancelled.push_back(&e);
            } else if (e.want_pause) {
                e.want_pause = false;
                if (e.state == ReqState::kRunning) edits.push_back({&e, 2});
                else if (e.state == ReqState::kQueued) e.state = ReqState::kPaused;
            } else if (e.want_resume) {
                e.want_resume = false;
                if (e.state == ReqState::kPaused) edits.push_back({&e, kReqActive});
            }
        }
        std::vector<SweepRequest*> running;
        lk.unlock();

        // 2. device work without the lock. Fresh requests are prepared together;
        // if that fails, one at a time, so that only those that fail alone (a
        // recount list that does not fit, a device error) fail.
        bool ok = true;
        std::vector<SweepRequest*> fresh;
        for (Entry* e : admit) fresh.push_back(&e->req);
        std::vector<Entry*> rejected;
        if (!fresh.empty() && !(sw_.prepare(fresh) && sw_.upload(fresh))) {
            (void)hipGetLastError();
            std::vector<Entry*> kept;
            for (Entry* e : admit) {
                const std::vector<SweepRequest*> one{&e->req};
                if (sw_.prepare(one) && sw_.upload(one)) {
                    kept.push_back(e);
                } else {
                    (void)hipGetLastError();
                    (void)sw_.set_status(e->req, 0);
                    rejected.push_back(e);
                }
            }
            admit.swap(kept);
        }
        for (Entry* e : admit) {
            // a pause that arrived while queued holds it right away
            if (ok && e->state == ReqState::kPaused) ok = sw_.set_status(e->req, 2);
        }
        for (auto& ed : edits) ok = ok && sw_.set_status(ed.first->req, ed.second);
        for (Entry* e : cancelled) ok = ok && sw_.set_status(e->req, 0);
        RoundStats rs{};
        lk.lock();
        for (Entry* e : rejected) {
            e->state = ReqState::kFailed;
            e->t_done = Clock::now();
            sw_.unreserve(e->req);
            in_use_ -= e->footprint;
            admit_blocked_ = false;
            admitted_.erase(std::find(admitted_.begin(), admitted_.end(), e->id));
        }
        for (Entry* e : admit)
            if (e->state == ReqState::kQueued) e->state = ReqState::kRunning;
        for (auto& ed : edits)
            ed.first->state = ed.second == kReqActive ? ReqState::kRunning : ReqState::kPaused;
        for (int64_t id : admitted_) {
            Entry& e = *entries_[id];
            if (e.state == ReqState::kRunning &&
                std::find(cancelled.begin(), cancelled.end(), &e) == cancelled.end())
                running.push_back(&e.req);
        }
        // With nothing queued and every running request more than a few
        // rounds from its end, no latency is at stake: rounds twice as long
        // (fewer launch tails); a new arrival brings short rounds back.
        bool long_rounds = queue_.empty();
        for (const SweepRequest* r : running)
            long_rounds = long_rounds && r->remaining_ns > 4.0 * (double)round_ns_;
        lk.unlock();

        // 3. the next round; the older round once a newer one is queued (or
        // nothing more was)
        int launched = 0;
        if (ok && !running.empty()) {
            RoundPolicy pol;
            pol.round_ns = long_rounds ? 2 * round_ns_ : round_ns_;
            pol.shortest_first = true;
            pol.direct = !persistent_;
            if (persistent_) {
                ok = sw_.round(running, pol, &rs);
            } else {
                launched = sw_.launch_round(running, pol);
                ok = launched >= 0;
            }
        }
        if (ok && sw_.in_flight() > 0 && (sw_.in_flight() > 1 || launched == 0))
            ok = sw_.wait_round(&rs);
        // a cancelled request is not in the round just launched; its earlier
        // rounds have ended before its partial result is ranked
        for (Entry* e : cancelled)
            while (ok && e->req.inflight > 0 && sw_.in_flight() > 0) ok = sw_.wait_round(&rs);
        lk.lock();
        finish(ok, cancelled, rs);
    }
    // rounds still in flight end before the session's device memory goes
    while (sw_.in_flight() > 0 && sw_.wait_round(nullptr)) {
    }
}

}  // namespace slime
