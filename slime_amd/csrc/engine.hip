// Host orchestration of one search: slabs along z, score histogram, emit
// gate, overflow retry and exact top-K ranking.
//
// Reference behaviour followed (SlimeCoreGPU.cu:3164-3915):
//   - the region is scanned in z slabs; after each slab the hit counter and
//     the buffer are read back; an overflowing slab is rerun (:3795-3805);
//   - every in-range hit is counted, but only hits that can still reach the
//     global top-K are materialised (emit gate, :3731-3740);
//   - ranking: size desc, center_x asc, center_z asc (:213-220, 3860-3880);
//   - pause is a sleep-poll between slabs, cancel ranks what was found
//     (:3754-3757).
// The gate here is exact and needs no completion pass: it is raised only to
// a score s for which at least K hits with score >= s were already counted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>

#include "engine.h"

namespace slime {

Control& control() {
    static Control c;
    return c;
}

namespace {

constexpr uint32_t kHitCap = 1u << 24;  // 16M hits (192 MB) per slab
constexpr int64_t kTargetCentersPerLaunch = int64_t(1) << 37;

struct Buffers {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    Hit* d_hits = nullptr;
    unsigned long long* d_counters = nullptr;  // [0] emitted, [1..222] histogram
    unsigned long long* h_counters = nullptr;  // pinned
    Hit* h_hits = nullptr;                     // pinned
    size_t h_hits_cap = 0;
};

Buffers g_buf;
std::mutex g_mutex;
std::atomic<int64_t> g_kernel_launches{0};
std::atomic<int64_t> g_kernel_ns{0};

void free_buffers() {
    if (g_buf.device >= 0) hipSetDevice(g_buf.device);
    if (g_buf.d_hits) hipFree(g_buf.d_hits);
    if (g_buf.d_counters) hipFree(g_buf.d_counters);
    if (g_buf.h_counters) hipHostFree(g_buf.h_counters);
    if (g_buf.h_hits) hipHostFree(g_buf.h_hits);
    if (g_buf.ev_start) hipEventDestroy(g_buf.ev_start);
    if (g_buf.ev_stop) hipEventDestroy(g_buf.ev_stop);
    if (g_buf.stream) hipStreamDestroy(g_buf.stream);
    g_buf = Buffers{};
}

bool ensure_buffers() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (g_buf.device == dev && g_buf.d_hits) return true;
    free_buffers();
    g_buf.device = dev;
    if (hipStreamCreateWithFlags(&g_buf.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&g_buf.ev_start) != hipSuccess || hipEventCreate(&g_buf.ev_stop) != hipSuccess ||
        hipMalloc(&g_buf.d_hits, sizeof(Hit) * (size_t)kHitCap) != hipSuccess ||
        hipMalloc(&g_buf.d_counters, sizeof(unsigned long long) * (1 + kHistBins)) != hipSuccess ||
        hipHostMalloc(&g_buf.h_counters, sizeof(unsigned long long) * (1 + kHistBins),
                      hipHostMallocDefault) != hipSuccess) {
        free_buffers();
        return false;
    }
    return true;
}

bool ensure_host_hits(size_t n) {
    if (g_buf.h_hits_cap >= n) return true;
    if (g_buf.h_hits) hipHostFree(g_buf.h_hits);
    g_buf.h_hits = nullptr;
    g_buf.h_hits_cap = 0;
    size_t want = std::max<size_t>(n, 1 << 16);
    if (hipHostMalloc(&g_buf.h_hits, sizeof(Hit) * want, hipHostMallocDefault) != hipSuccess) {
        g_buf.h_hits = nullptr;
        return false;
    }
    g_buf.h_hits_cap = want;
    return true;
}

// Largest score s with at least k counted hits of score >= s; `floor` if none.
int32_t gate_for(const int64_t* hist, int64_t k, int32_t floor) {
    int64_t acc = 0;
    for (int32_t s = kHistBins - 1; s >= 0; --s) {
        acc += hist[s];
        if (acc >= k) return std::max(s, floor);
    }
    return floor;
}

// Keeps only the k best candidates (unordered).
void trim(std::vector<Hit>& cand, size_t k) {
    if (cand.size() <= k) return;
    std::nth_element(cand.begin(), cand.begin() + k, cand.end(), ranks_before);
    cand.resize(k);
}

struct Geometry {
    int32_t tiles_x;
    int32_t tile_h;
    int64_t slab_h;
};

Geometry plan(int64_t w, int64_t h) {
    Geometry g;
    g.tiles_x = (int32_t)((w + kTileOut - 1) / kTileOut);
    // Aim for >= 2048 workgroups per launch; tall tiles amortise the 16-row halo.
    int64_t want_h = (h * g.tiles_x + 2047) / 2048;
    int64_t th = std::min<int64_t>(2048, std::max<int64_t>(16, (want_h + 3) / 4 * 4));
    g.tile_h = (int32_t)th;
    int64_t slab = kTargetCentersPerLaunch / std::max<int64_t>(1, w);
    slab = std::max<int64_t>(th, slab / th * th);
    // keep the grid of one launch within int32 workgroups
    int64_t max_tiles_z = (int64_t(1) << 30) / g.tiles_x;
    slab = std::min<int64_t>(slab, std::max<int64_t>(1, max_tiles_z) * th);
    g.slab_h = slab;
    return g;
}

void wait_while_paused() {
    Control& c = control();
    while (c.pause.load(std::memory_order_relaxed) && !c.cancel.load(std::memory_order_relaxed))
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
}

}  // namespace

void kernel_stats(int64_t* launches, int64_t* kernel_ns, bool reset) {
    if (launches) *launches = g_kernel_launches.load();
    if (kernel_ns) *kernel_ns = g_kernel_ns.load();
    if (reset) {
        g_kernel_launches.store(0);
        g_kernel_ns.store(0);
    }
}

void release_engine() {
    std::lock_guard<std::mutex> lock(g_mutex);
    free_buffers();
}

int64_t run_search(const SearchSpec& spec, ResultRow* out, int32_t max_results,
                   int64_t* hist_out) {
    std::lock_guard<std::mutex> lock(g_mutex);
    Control& ctl = control();
    if (!ensure_buffers()) {
        hipGetLastError();
        return -1;
    }
    const int64_t W = spec.w, H = spec.h;
    const Geometry geo = plan(W, H);
    int64_t hist[kHistBins] = {0};
    std::vector<Hit> cand;
    const size_t keep = (size_t)max_results;
    int32_t emit_min = spec.min_size;
    int64_t slab_h = geo.slab_h;

    for (int64_t z_off = 0; z_off < H;) {
        wait_while_paused();
        if (ctl.cancel.load(std::memory_order_relaxed)) break;
        const int64_t rows = std::min<int64_t>(slab_h, H - z_off);
        ScanParams p{};
        p.seed = spec.seed;
        p.x0 = spec.x0;
        p.z0 = (int32_t)((uint32_t)spec.z0 + (uint32_t)z_off);
        p.w = (int32_t)W;
        p.h = (int32_t)rows;
        p.tile_h = geo.tile_h;
        p.tiles_x = geo.tiles_x;
        p.tiles_total = (int32_t)(geo.tiles_x * ((rows + geo.tile_h - 1) / geo.tile_h));
        p.min_size = spec.min_size;
        p.max_size = spec.max_size;
        p.emit_min = emit_min;
        p.use_annulus = spec.rd_min > 0;
        p.rd_min_sq = spec.rd_min * spec.rd_min;
        p.acx = spec.acx;
        p.acz = spec.acz;
        p.hits = g_buf.d_hits;
        p.hit_cap = kHitCap;
        p.emitted = g_buf.d_counters;
        p.hist = g_buf.d_counters + 1;

        const auto t0 = std::chrono::steady_clock::now();
        if (hipMemsetAsync(g_buf.d_counters, 0, sizeof(unsigned long long) * (1 + kHistBins),
                           g_buf.stream) != hipSuccess ||
            hipEventRecord(g_buf.ev_start, g_buf.stream) != hipSuccess ||
            launch_scan_plain(p, g_buf.stream) != hipSuccess ||
            hipEventRecord(g_buf.ev_stop, g_buf.stream) != hipSuccess ||
            hipMemcpyAsync(g_buf.h_counters, g_buf.d_counters,
                           sizeof(unsigned long long) * (1 + kHistBins), hipMemcpyDeviceToHost,
                           g_buf.stream) != hipSuccess ||
            hipStreamSynchronize(g_buf.stream) != hipSuccess) {
            hipGetLastError();
            ctl.progress.store(0, std::memory_order_relaxed);
            return -1;
        }
        const auto t1 = std::chrono::steady_clock::now();
        float kernel_ms = 0.f;
        if (hipEventElapsedTime(&kernel_ms, g_buf.ev_start, g_buf.ev_stop) == hipSuccess) {
            g_kernel_launches.fetch_add(1, std::memory_order_relaxed);
            g_kernel_ns.fetch_add((int64_t)((double)kernel_ms * 1e6), std::memory_order_relaxed);
        }
        const unsigned long long emitted = g_buf.h_counters[0];
        const unsigned long long* slab_hist = g_buf.h_counters + 1;

        if (emitted > kHitCap) {
            // Raise the gate using this slab's counts, or split the slab.
            int64_t merged[kHistBins];
            for (int s = 0; s < kHistBins; ++s) merged[s] = hist[s] + (int64_t)slab_hist[s];
            const int32_t next = gate_for(merged, (int64_t)keep, spec.min_size);
            if (next > emit_min) {
                emit_min = next;
                continue;
            }
            if (rows > geo.tile_h) {
                slab_h = std::max<int64_t>(geo.tile_h, rows / 2 / geo.tile_h * geo.tile_h);
                continue;
            }
            ctl.progress.store(0, std::memory_order_relaxed);
            return -1;
        }
        for (int s = 0; s < kHistBins; ++s) hist[s] += (int64_t)slab_hist[s];
        if (emitted > 0) {
            if (!ensure_host_hits(emitted) ||
                hipMemcpyAsync(g_buf.h_hits, g_buf.d_hits, sizeof(Hit) * emitted,
                               hipMemcpyDeviceToHost, g_buf.stream) != hipSuccess ||
                hipStreamSynchronize(g_buf.stream) != hipSuccess) {
                hipGetLastError();
                ctl.progress.store(0, std::memory_order_relaxed);
                return -1;
            }
            cand.insert(cand.end(), g_buf.h_hits, g_buf.h_hits + emitted);
            if (cand.size() > std::max<size_t>(2 * keep, 1 << 20)) trim(cand, keep);
        }
        emit_min = std::max(emit_min, gate_for(hist, (int64_t)keep, spec.min_size));
        ctl.processed_centers.fetch_add(rows * W, std::memory_order_relaxed);
        ctl.scan_work_ns.fetch_add(
            std::max<int64_t>(1, std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count()),
            std::memory_order_relaxed);
        z_off += rows;
        ctl.progress.store(1 + (int32_t)((z_off * 79) / std::max<int64_t>(1, H)),
                           std::memory_order_relaxed);
    }

    int64_t total = 0;
    for (int s = 0; s < kHistBins; ++s) total += hist[s];
    if (hist_out) std::memcpy(hist_out, hist, sizeof(hist));
    const size_t n = std::min(cand.size(), keep);
    if (n < cand.size())
        std::partial_sort(cand.begin(), cand.begin() + n, cand.end(), ranks_before);
    else
        std::sort(cand.begin(), cand.end(), ranks_before);
    for (size_t i = 0; i < n; ++i) {
        const int64_t bx = (int64_t)cand[i].cx * 16 + 8;
        const int64_t bz = (int64_t)cand[i].cz * 16 + 8;
        out[i].size = cand[i].size;
        out[i].center_x = (int32_t)(uint32_t)(uint64_t)bx;
        out[i].center_z = (int32_t)(uint32_t)(uint64_t)bz;
        out[i].obs_count = 0;
        out[i].afk_x = out[i].center_x;
        out[i].afk_z = out[i].center_z;
    }
    ctl.progress.store(90, std::memory_order_relaxed);
    return total;
}

}  // namespace slime
