// Multi-search sweep: device-resident state of many searches scanned by the
// persistent per-class kernels (batched calls and the served session).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"
#include "search_util.h"

namespace slime {

// Measured cost of each kernel class: device ns per tile row at full
// occupancy (the whole GPU busy with that class) and a fixed per-tile cost
// in rows.
struct CostTable {
    bool ready = false;
    double ns_row[kSweepClasses] = {};
    double extra_rows[kSweepClasses] = {};
};

// Host state of one search in the sweeper.
struct SweepRequest {
    SearchSpec spec;
    int32_t keep = 0;  // max_results
    // plan
    std::vector<BandJob> jobs;
    std::vector<int32_t> job_cls;  // kernel index of each job (sweep_kernel_of)
    int64_t centers = 0;
    double est_ns = 0;   // modelled device time of the whole request
    int64_t hit_cap = 0;
    // reservation
    int32_t slot = -1;
    int64_t job_off = -1, hit_off = -1;
    // progress
    int32_t status = 0;  // mirror of the device status word
    std::vector<uint32_t> done;     // per job: tiles finished (the first done[j] of the job)
    std::vector<uint32_t> planned;  // per job: tiles launched (direct rounds)
    int32_t inflight = 0;           // launched rounds holding tiles of it
    double remaining_ns = 0;        // modelled device time of its tiles not yet launched
    bool complete = false;
    // first-draw fixes
    MismatchIndex mm;
    std::vector<CenterFix> fixes;
    // result
    int64_t total = -1;
    bool fallback = false;
    bool truncated = false;  // cancelled with more hits than its region held
    int64_t hits_stored = 0;  // hits in the request's device region at the end
    int64_t hits_copied = 0;  // of them, brought to the host to rank
    std::vector<ResultRow> rows;
};

struct RoundPolicy {
    int64_t round_ns = 0;         // device time per round, over its class launches (0: unbounded)
    bool shortest_first = false;  // claim order: least work left first (else largest tiles first)
    bool use_graph = false;       // replay the launch sequence from a captured hipGraph
    // one workgroup per tile over a range of tiles per job, no claims (the
    // requests' status words are not read: pause and cancel act between rounds)
    bool direct = false;
};

struct RoundStats {
    int64_t kernel_ns = 0;
    int32_t launches = 0;
    int32_t rounds = 0;
};

class Sweeper {
   public:
    ~Sweeper();
    bool init(int32_t req_cap, int64_t job_cap, int64_t hit_cap);
    void release();
    bool ready() const { return device_ >= 0; }
    int32_t req_cap() const { return req_cap_; }
    int64_t job_cap() const { return job_cap_; }
    int64_t hit_cap() const { return hit_cap_; }
    int32_t free_slots() const { return (int32_t)free_slots_.size(); }
    const CostTable& costs() const { return costs_; }
    void set_costs(const CostTable& c) { costs_ = c; }

    bool calibrate();                         // fills the cost table (once)
    double class_tile_ns(int cls, int32_t tile_h) const;  // measured ns of a tile of a class
    double tile_ns(int kernel, int32_t tile_h) const;     // of a tile of a kernel
    bool plan(SweepRequest& r) const;         // schedule + hit capacity
    static int64_t footprint(const SweepRequest& r);  // device bytes it holds
    bool reserve(SweepRequest& r);            // slot, job range, hit region
    void unreserve(SweepRequest& r);
    bool prepare(const std::vector<SweepRequest*>& reqs);  // mismatch fixes
    bool upload(const std::vector<SweepRequest*>& reqs);   // jobs, cleared outputs, active
    bool set_status(SweepRequest& r, int32_t status);
    bool clear_status();  // every request slot inactive
    // One round, waited for (after any rounds still in flight).
    bool round(const std::vector<SweepRequest*>& reqs, const RoundPolicy& pol, RoundStats* st);
    // A round launched without waiting for it, planned from the tiles already
    // launched (direct rounds: up to two in flight, the next one queued while
    // the last one runs): 1 launched, 0 nothing to launch, -1 error.
    int launch_round(const std::vector<SweepRequest*>& reqs, const RoundPolicy& pol);
    // Waits for the oldest round in flight and takes in its done counts; its
    // requests with every tile done are complete.
    bool wait_round(RoundStats* st);
    int in_flight() const { return pend_n_; }
    // whether the oldest round in flight has ended (its done counts copied)
    bool oldest_ended() const {
        return pend_n_ > 0 && hipEventQuery(ev_copied_[pend_first_]) == hipSuccess;
    }
    // host time by which the newest round in flight is modelled to end
    std::chrono::steady_clock::time_point expected_end() const {
        return pend_[(pend_first_ + pend_n_ + kRing - 1) % kRing].t_end;
    }
    double progress(const SweepRequest& r) const;
    bool finalize(SweepRequest& r, bool partial);
    // Many finished requests: their counts in one copy, their stored hits in
    // one more round trip (device top-K selections one request at a time).
    bool finalize_all(const std::vector<SweepRequest*>& rs, bool partial);

   private:
    bool ensure_fix_cap(size_t n);
    bool prepare_fixes(const std::vector<SweepRequest*>& reqs);
    // Ranking of a finished request, in three steps: the corrected histogram,
    // total and threshold from its counts (or the rerun alone of an
    // overflowed region); its stored hits (or the device selection of those
    // that can reach the top K); the ranked rows.
    struct RankState {
        bool done = false;  // rerun alone: rows final
        size_t stored = 0;
        int32_t t = 0;
        bool select = false;
        std::vector<Hit> cand;
    };
    bool rank_begin(SweepRequest& r, bool partial, const unsigned long long* hist,
                    unsigned long long emitted, RankState& rs);
    bool rank_hits(SweepRequest& r, const RankState& rs, std::vector<Hit>& hits);
    void rank_end(SweepRequest& r, RankState& rs, const std::vector<Hit>& hits);

    int device_ = -1;
    int cus_ = 256;
    int grid_[kSweepKernels] = {};         // persistent launches: resident workgroups
    int direct_grid_[kSweepKernels] = {};  // one-tile launches: resident workgroups
    hipStream_t stream_ = nullptr;  // set-up, uploads, status words, ranking
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    // Rounds run on the lane streams (a round's class launches side by side)
    // from a ring of kRing slots, each with its own launch arrays, done-count
    // copy and events; at most kRing - 1 rounds are in flight, so a slot is
    // rewritten only after its round has been waited for and the round before
    // the one being waited for keeps its events (the device time overlap).
    static constexpr int kLanes = 3;
    static constexpr int kRing = 3;
    hipStream_t lane_[kLanes] = {};
    struct Expect {
        SweepRequest* req;
        int32_t j;
        uint32_t count;  // the job's done count once the round has ended
    };
    struct Pending {
        bool direct = true;
        int32_t launches = 0;
        double est_ns = 0;
        std::chrono::steady_clock::time_point t_launch, t_end;
        std::vector<int32_t> order, starts;  // host copies of the launch arrays
        std::vector<int64_t> firsts;
        std::vector<Expect> expect;
        std::vector<SweepRequest*> reqs;
    };
    Pending pend_[kRing];
    int pend_first_ = 0, pend_n_ = 0, last_slot_ = -1;
    int big_lane_ = 0;  // lane of the last round's longest launch
    hipEvent_t rev0_[kRing] = {}, rev1_[kRing] = {}, ev_copied_[kRing] = {}, ev_fork_[kRing] = {};
    hipEvent_t ev_join_[kRing][kLanes - 1] = {};
    int32_t req_cap_ = 0;
    int64_t job_cap_ = 0, hit_cap_ = 0, fix_cap_ = 0;
    SweepJob* d_jobs_ = nullptr;
    uint32_t* d_cursor_ = nullptr;
    uint32_t* d_done_ = nullptr;
    uint32_t* h_done_[kRing] = {};
    int32_t* d_order_[kRing] = {};
    int64_t* d_first_[kRing] = {};  // direct launches: first block of each job (job_cap + kernels)
    int32_t* d_start_[kRing] = {};  // direct launches: first tile of each job
    int32_t* d_status_ = nullptr;
    unsigned long long* d_hist_ = nullptr;
    unsigned long long* d_emitted_ = nullptr;
    Hit* d_hits_ = nullptr;
    uint32_t* d_claims_ = nullptr;
    uint64_t* d_table_ = nullptr;
    int32_t table_n_ = 0;
    MismatchQuery* d_queries_ = nullptr;
    TaggedMismatch* d_mm_ = nullptr;
    unsigned int* d_mm_count_ = nullptr;
    CenterFix* d_fixes_ = nullptr;
    Hit* d_select_ = nullptr;  // a finishing request's hits that can still rank (kSelectCap)
    unsigned int* d_select_count_ = nullptr;
    CostTable costs_;
    std::vector<int32_t> free_slots_;
    struct Ranges {
        std::map<int64_t, int64_t> free_;
        void reset(int64_t n) {
            free_.clear();
            if (n > 0) free_[0] = n;
        }
        int64_t alloc(int64_t n);
        void release(int64_t off, int64_t n);
    };
    Ranges jobs_alloc_, hits_alloc_;
    std::map<std::vector<int64_t>, hipGraphExec_t> graphs_;
};

// ---- batched searches ----

struct SweepOptions {
    bool use_graph = false;
    // the persistent kernels fed from the device work queue (per-job claim
    // cursors) instead of one workgroup per tile
    bool persistent = false;
};

struct BatchStats {
    int64_t centers = 0;    // centers decided
    int64_t kernel_ns = 0;  // device time of the sweep launches
    int64_t prepare_ns = 0; // planning, mismatch fixes, upload (host clock)
    int64_t sweep_ns = 0;   // the sweep launches incl. sync (host clock)
    int64_t wall_ns = 0;
    int32_t launches = 0;
    int32_t jobs = 0;
    int32_t fallbacks = 0;  // searches rerun alone after a hit-region overflow
    int64_t hits_stored = 0;  // hits in the searches' device regions
    int64_t hits_copied = 0;  // of them, brought to the host to rank (top-K selection)
    int64_t class_centers[kSweepClasses] = {};  // centers decided per kernel class
};

// Runs n searches in one sweep. out holds n * max_results rows (search i at
// i * max_results); totals[i] = its hit count. Returns 0, or -1 on a device
// error.
int run_batch(const SearchSpec* specs, int32_t n, int32_t max_results, ResultRow* out,
              int64_t* totals, const SweepOptions& opt, BatchStats* stats);
void last_batch_stats(BatchStats* st);
// The autotuner's measured table (kSweepClasses entries each).
bool sweep_cost_table(double* ns_row, double* extra_rows);

}  // namespace slime

namespace slime {

// ---- served searches: continuous batching --------------------------------------
//
// A session owns a sweeper and a worker thread. Requests are admitted in
// arrival order while their device footprint fits the HBM budget and a slot
// is free; every round the worker applies pause/resume/cancel as writes of
// the requests' device status words, admits what fits, runs one bounded
// launch per kernel class over the tiles of all active requests (least work
// left first; each job's done count stays on the device) and finishes the
// requests whose tiles are all done.
enum class ReqState : int32_t {
    kQueued = 0,
    kRunning = 1,
    kPaused = 2,
    kDone = 3,
    kCancelled = 4,
    kFailed = 5,
};

struct SessionStats {
    int64_t centers_done = 0;  // centers of completed requests
    int64_t kernel_ns = 0;
    int32_t rounds = 0;
    int32_t active = 0, queued = 0;
    int64_t hbm_in_use = 0;
    int64_t hbm_budget = 0;
};

class Session {
   public:
    // persistent: rounds of the persistent work-queue kernels, which read each
    // request's status word before every claim (pause and cancel act within a
    // launch); otherwise direct rounds, one workgroup per tile over a range of
    // tiles per job, with the per-search kernels' register budget (pause and
    // cancel act between rounds).
    Session(int64_t hbm_budget, int32_t max_active, int32_t round_ms, bool persistent = false);
    ~Session();
    bool ok() const { return ok_; }
    int64_t submit(const SearchSpec& spec, int32_t keep);
    bool pause(int64_t id);
    bool resume(int64_t id);
    bool cancel(int64_t id);
    // state, progress in [0, 1], total (-1 until done), latency submit -> done
    bool poll(int64_t id, ReqState* state, double* progress, int64_t* total, int64_t* latency_ns);
    ReqState wait(int64_t id, int32_t timeout_ms);
    int64_t result(int64_t id, ResultRow* out, int32_t cap);
    bool release(int64_t id);
    SessionStats stats();

   private:
    struct Entry;
    void worker();
    bool ok_ = false;
    int64_t budget_ = 0;
    int32_t max_active_ = 0;
    int64_t round_ns_ = 0;
    bool persistent_ = false;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::thread thread_;
    bool stop_ = false;
    Sweeper sw_;
    std::map<int64_t, std::unique_ptr<Entry>> entries_;
    std::deque<int64_t> queue_;
    std::vector<int64_t> admitted_;
    int64_t next_id_ = 1;
    int64_t in_use_ = 0;
    bool admit_blocked_ = false;  // the sweeper had no contiguous room for the queue head
    SessionStats st_;
};

}  // namespace slime
