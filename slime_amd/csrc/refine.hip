// Precise AFK refinement kernel: one workgroup per result row.
//
// The workgroup tests the 23 x 23 chunks around the row's center once
// (floor(center / 16) - 11 .. + 11: every block within 128 of any of the 81
// AFK points lies in it) and keeps them as 23 z-bitmasks in LDS. A point
// (ox, oz) at height difference dy then counts, for each block column
// x = ox + dx with dx^2 + dy^2 <= 16384, the slime blocks of that column in
// the ring's z intervals [oz - ro, oz - ri - 1] and [oz + ri + 1, oz + ro]
// (ro, ri the outer and inner half-widths). A slime-block prefix count F(z)
// of a chunk column needs one popcount of its mask below z's chunk, so an
// interval costs two F's; the ring's half-widths per (dy, dx) come from an
// LDS table built per group of heights. The outer disk and the inner disk
// (|dx| <= 24 only) are separate loops. This is the set the reference counts
// per slime chunk and column (SlimeCore.cpp:234-331); no weight tables and
// no per-point chunk caches are needed.
//
// Work item = (point, height); lanes of a wave share the height, so the ring
// table reads are broadcasts. Each thread keeps its best (count, order) and
// the workgroup reduces them with wave shuffles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "refine.h"
#include "slime_math.h"

namespace slime {
namespace {

constexpr int kRefineThreads = 256;
constexpr int kWin = 23;             // chunks per side of the window
constexpr int kSpan = 257;           // dx = -128..128
constexpr int kMaxGroup = 8;         // heights per LDS ring table
constexpr int kOffsets = 81;         // 9 x 9 AFK points, step 4 blocks
constexpr int kLevelBits = 9;        // active heights <= 257 < 512

struct RefineArgs {
    ResultRow* rows;
    const int64_t* row_seeds;  // per-row world seed, or null for `seed`
    int64_t seed;
    int32_t noff;        // kOffsets or 1
    int32_t platform_y;
    int32_t lv_n;        // active heights (|platform_y - y| <= 128)
    int32_t y_first;     // first active height
    int32_t y_step;      // between active heights
    int32_t y_level0;    // height of level 0 (all-zero answer of kRuleFirstOrder)
    int32_t group;       // heights per ring table
    int32_t rule;
    int32_t pack;
};

// floor(sqrt(v)) for 0 <= v <= 16384.
__device__ __forceinline__ int32_t isqrt_small(int32_t v) {
    int32_t r = (int32_t)sqrtf((float)v);
    while ((r + 1) * (r + 1) <= v) ++r;
    while (r * r > v) --r;
    return r;
}

// Ring half-widths of column dx at height difference squared dy2: byte 0 =
// ro = floor(sqrt(16384 - dy2 - dx^2)), byte 1 = ri = floor(sqrt(576 - dy2
// - dx^2)) where that radicand is >= 0. Columns outside the outer disk are
// never read (the loops stop at the disk's reach).
__device__ __forceinline__ uint32_t ring_entry(int32_t dy2, int32_t dx) {
    const int32_t outer2 = 16384 - dy2 - dx * dx;
    if (outer2 < 0) return 0;
    const int32_t inner2 = 576 - dy2 - dx * dx;
    const uint32_t ri = inner2 >= 0 ? (uint32_t)isqrt_small(inner2) : 0u;
    return (uint32_t)isqrt_small(outer2) | (ri << 8);
}

// Slime blocks of one block column in window z range [lo, hi] (mask bit j =
// chunk j of the column): F(hi + 1) - F(lo) with F(z) = 16 * (slime chunks
// below z's chunk) + (z's chunk is slime) * (z mod 16).
__device__ __forceinline__ int32_t column_count(uint32_t mask, uint32_t lo, uint32_t hi1) {
    const uint32_t jl = lo >> 4, jh = hi1 >> 4;
    const uint32_t below = __popc(__builtin_amdgcn_ubfe(mask, 0, jh)) -
                           __popc(__builtin_amdgcn_ubfe(mask, 0, jl));
    return (int32_t)((below << 4) + __umul24(__builtin_amdgcn_ubfe(mask, jh, 1), hi1 & 15u) -
                     __umul24(__builtin_amdgcn_ubfe(mask, jl, 1), lo & 15u));
}

__device__ __forceinline__ bool better(int32_t obs, int32_t key, int32_t bobs, int32_t bkey) {
    return obs > bobs || (obs == bobs && key < bkey);
}

__device__ __forceinline__ void offset_of(int32_t noff, int32_t k, int32_t& dx, int32_t& dz) {
    dx = noff > 1 ? -16 + (k / 9) * 4 : 0;
    dz = noff > 1 ? -16 + (k % 9) * 4 : 0;
}

__global__ __launch_bounds__(kRefineThreads) void refine_kernel(RefineArgs a) {
    __shared__ uint32_t win[kWin];
    __shared__ uint32_t ring[kMaxGroup * kSpan];
    __shared__ int32_t reach[kMaxGroup];     // outer disk: |dx| <= reach
    __shared__ int32_t reach_in[kMaxGroup];  // inner disk: |dx| <= reach_in (-1: none)
    __shared__ int32_t red_obs[kRefineThreads / 64];
    __shared__ int32_t red_key[kRefineThreads / 64];

    const int tid = threadIdx.x;
    ResultRow& row = a.rows[blockIdx.x];
    const int64_t seed = a.row_seeds ? a.row_seeds[blockIdx.x] : a.seed;
    const int32_t bx = row.center_x, bz = row.center_z;
    const int32_t ubx = (bx >> 4) - 11, ubz = (bz >> 4) - 11;

    if (tid < kWin) win[tid] = 0;
    __syncthreads();
    for (int c = tid; c < kWin * kWin; c += kRefineThreads) {
        const int i = c / kWin, j = c - i * kWin;
        if (is_slime(seed, ubx + i, ubz + j)) atomicOr(&win[i], 1u << j);
    }

    // Center relative to the window's first block (160..207 after the offset).
    const int32_t xb = (bx & 15) + 176, zb = (bz & 15) + 176;
    int32_t bobs = -1, bkey = INT_MAX;
    for (int g0 = 0; g0 < a.lv_n; g0 += a.group) {
        const int gn = min(a.group, a.lv_n - g0);
        __syncthreads();  // window built / previous table no longer read
        for (int e = tid; e < gn * kSpan; e += kRefineThreads) {
            const int l = e / kSpan, dx = e - l * kSpan - 128;
            const int32_t dy = a.platform_y - (a.y_first + (g0 + l) * a.y_step);
            ring[e] = ring_entry(dy * dy, dx);
            if (dx == 0) {
                reach[l] = isqrt_small(16384 - dy * dy);
                reach_in[l] = dy * dy <= 576 ? isqrt_small(576 - dy * dy) : -1;
            }
        }
        __syncthreads();
        const int items = gn * a.noff;
        for (int t = tid; t < items; t += kRefineThreads) {
            const int l = t / a.noff, k = t - l * a.noff;
            int32_t dxo, dzo;
            offset_of(a.noff, k, dxo, dzo);
            const int32_t x0 = xb + dxo, z0 = zb + dzo;
            const uint32_t* rt = ring + l * kSpan + 128;
            // Outer disk z in [z0 - ro, z0 + ro], then minus the inner disk.
            int32_t obs = 0;
            for (int dx = -reach[l]; dx <= reach[l]; ++dx) {
                const uint32_t ro = rt[dx] & 0xFFu;
                obs += column_count(win[(uint32_t)(x0 + dx) >> 4], (uint32_t)z0 - ro,
                                    (uint32_t)z0 + ro + 1u);
            }
            for (int dx = -reach_in[l]; dx <= reach_in[l]; ++dx) {
                const uint32_t ri = (rt[dx] >> 8) & 0xFFu;
                obs -= column_count(win[(uint32_t)(x0 + dx) >> 4], (uint32_t)z0 - ri,
                                    (uint32_t)z0 + ri + 1u);
            }
            const int32_t key = (k << kLevelBits) | (g0 + l);
            if (better(obs, key, bobs, bkey)) {
                bobs = obs;
                bkey = key;
            }
        }
    }

    for (int s = 32; s >= 1; s >>= 1) {
        const int32_t o = __shfl_xor(bobs, s);
        const int32_t k = __shfl_xor(bkey, s);
        if (better(o, k, bobs, bkey)) {
            bobs = o;
            bkey = k;
        }
    }
    if ((tid & 63) == 0) {
        red_obs[tid >> 6] = bobs;
        red_key[tid >> 6] = bkey;
    }
    __syncthreads();
    if (tid != 0) return;
    for (int w = 1; w < kRefineThreads / 64; ++w) {
        if (better(red_obs[w], red_key[w], bobs, bkey)) {
            bobs = red_obs[w];
            bkey = red_key[w];
        }
    }
    int32_t obs = 0, dx = 0, dz = 0, y = a.platform_y;
    if (bobs > 0) {
        obs = bobs;
        offset_of(a.noff, bkey >> kLevelBits, dx, dz);
        y = a.y_first + ((bkey & ((1 << kLevelBits) - 1))) * a.y_step;
    } else if (a.rule == kRuleFirstOrder) {
        offset_of(a.noff, 0, dx, dz);
        y = a.y_level0;
    }
    row.obs_count = a.pack ? (int32_t)((uint32_t)obs | ((uint32_t)(y + 1024) << 20)) : obs;
    row.afk_x = (int32_t)((uint32_t)bx + (uint32_t)dx);
    row.afk_z = (int32_t)((uint32_t)bz + (uint32_t)dz);
}

struct RefineState {
    std::mutex mu;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ResultRow* d_rows = nullptr;
    int64_t* d_seeds = nullptr;
    size_t cap = 0;
    int64_t last_ns = 0, last_evals = 0;
};

RefineState& state() {
    static RefineState s;
    return s;
}

int64_t floor_div(int64_t a, int64_t b) { return a >= 0 ? a / b : -((b - 1 - a) / b); }

}  // namespace

YLevels y_levels(int32_t enabled, int32_t platform_y, int32_t y_min, int32_t y_max,
                 int32_t y_step) {
    YLevels L;
    L.platform_y = platform_y;
    int64_t start = enabled ? y_min : platform_y;
    int64_t end = enabled ? y_max : platform_y;
    int64_t step = enabled ? y_step : 1;
    if (step <= 0) step = 1;
    if (end < start) std::swap(start, end);
    L.y_start = start;
    L.y_step = step;
    L.y_count = std::max<int64_t>(1, (end - start) / step + 1);
    return L;
}

int refine_rows(const RefineSpec& spec, ResultRow* rows, int32_t count) {
    return refine_rows_multi(spec, nullptr, rows, count);
}

namespace {

// The kernel's arguments for one refinement (rows and row_seeds left unset).
RefineArgs refine_args(const RefineSpec& spec) {
    const YLevels& L = spec.levels;
    // Heights within 128 blocks of the platform; the others count 0.
    const int64_t lo = std::max<int64_t>(
        0, -floor_div(-((int64_t)L.platform_y - 128 - L.y_start), L.y_step));
    const int64_t hi = std::min<int64_t>(
        L.y_count - 1, floor_div((int64_t)L.platform_y + 128 - L.y_start, L.y_step));
    const int32_t lv_n = hi >= lo ? (int32_t)(hi - lo + 1) : 0;

    RefineArgs a;
    a.seed = spec.seed;
    a.noff = spec.scan_xz ? kOffsets : 1;
    a.platform_y = L.platform_y;
    a.lv_n = lv_n;
    a.y_first = lv_n > 0 ? (int32_t)(L.y_start + lo * L.y_step) : L.platform_y;
    a.y_step = lv_n > 1 ? (int32_t)L.y_step : 0;
    a.y_level0 = (int32_t)L.y_start;
    a.group = a.noff == kOffsets ? kRefineThreads / kOffsets : kMaxGroup;
    a.rule = spec.rule;
    a.pack = spec.pack;
    return a;
}

}  // namespace

int refine_rows_multi(const RefineSpec& spec, const int64_t* row_seeds, ResultRow* rows,
                      int32_t count) {
    if (count <= 0) return 0;
    if (!rows) return -1;
    RefineArgs a = refine_args(spec);

    RefineState& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    if (!s.stream) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s.ev0) != hipSuccess || hipEventCreate(&s.ev1) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
    }
    if ((size_t)count > s.cap) {
        if (s.d_rows) (void)hipFree(s.d_rows);
        if (s.d_seeds) (void)hipFree(s.d_seeds);
        s.d_rows = nullptr;
        s.d_seeds = nullptr;
        s.cap = 0;
        const size_t cap = std::max<size_t>((size_t)count, 1024);
        if (hipMalloc(&s.d_rows, cap * sizeof(ResultRow)) != hipSuccess ||
            hipMalloc(&s.d_seeds, cap * sizeof(int64_t)) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        s.cap = cap;
    }
    a.rows = s.d_rows;
    a.row_seeds = row_seeds ? s.d_seeds : nullptr;
    const size_t bytes = (size_t)count * sizeof(ResultRow);
    if ((row_seeds && hipMemcpyAsync(s.d_seeds, row_seeds, (size_t)count * sizeof(int64_t),
                                     hipMemcpyHostToDevice, s.stream) != hipSuccess) ||
        hipMemcpyAsync(s.d_rows, rows, bytes, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
        hipEventRecord(s.ev0, s.stream) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)count), dim3(kRefineThreads), 0, s.stream,
                       a);
    // Results go to a host copy first so that a failed launch leaves rows as
    // they were.
    std::vector<ResultRow> out((size_t)count);
    if (hipGetLastError() != hipSuccess || hipEventRecord(s.ev1, s.stream) != hipSuccess ||
        hipMemcpyAsync(out.data(), s.d_rows, bytes, hipMemcpyDeviceToHost, s.stream) !=
            hipSuccess ||
        hipStreamSynchronize(s.stream) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s.ev0, s.ev1) == hipSuccess) s.last_ns = (int64_t)(ms * 1e6);
    s.last_evals = (int64_t)count * a.noff * a.lv_n;
    std::copy(out.begin(), out.end(), rows);
    return 0;
}

void refine_last_stats(int64_t* kernel_ns, int64_t* evaluations) {
    RefineState& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    if (kernel_ns) *kernel_ns = s.last_ns;
    if (evaluations) *evaluations = s.last_evals;
}

// ---- refinement without allocation ----------------------------------------
//
// The rows are staged in `scratch`, a file-scope device array of
// kNoAllocMaxRows rows (1.5 MiB; every device the library runs on holds its
// own copy until the process ends), and copied back through a static host
// array of the same size. Host code reaches the device array only through
// the address hipGetSymbolAddress gives for the current device, taken once per
// device at its first call (with the device current) together with a stream
// and two events. Nothing is allocated afterwards. Calls in one process are
// serialized by a lock; the array keeps the last call's rows.

namespace {

static __device__ __align__(256) ResultRow scratch[kNoAllocMaxRows];

constexpr int kNoAllocDevices = 64;

struct NoAllocDevice {
    bool ready = false;
    ResultRow* rows = nullptr;  // this device's `scratch`
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocDevices];
ResultRow g_noalloc_host[kNoAllocMaxRows];  // results before they replace the caller's rows

// Set-up of the current device (once): the array's address, a stream, events.
NoAllocDevice* noalloc_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) return nullptr;
    NoAllocDevice& d = g_noalloc[dev];
    if (d.ready) return &d;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    if (!d.stream && hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess)
        return nullptr;
    if (!d.ev0 && hipEventCreate(&d.ev0) != hipSuccess) return nullptr;
    if (!d.ev1 && hipEventCreate(&d.ev1) != hipSuccess) return nullptr;
    d.rows = static_cast<ResultRow*>(ptr);
    d.ready = true;
    return &d;
}

}  // namespace

int refine_rows_noalloc(const RefineSpec& spec, ResultRow* rows, int32_t count) {
    if (count < 0 || count > kNoAllocMaxRows) return -2;
    if (count == 0) return 0;
    if (!rows) return -1;
    RefineArgs a = refine_args(spec);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice* d = noalloc_device();
    if (!d) {
        (void)hipGetLastError();
        return -1;
    }
    a.rows = d->rows;
    a.row_seeds = nullptr;
    const size_t bytes = (size_t)count * sizeof(ResultRow);
    if (hipMemcpyAsync(d->rows, rows, bytes, hipMemcpyHostToDevice, d->stream) != hipSuccess ||
        hipEventRecord(d->ev0, d->stream) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)count), dim3(kRefineThreads), 0, d->stream,
                       a);
    if (hipGetLastError() != hipSuccess || hipEventRecord(d->ev1, d->stream) != hipSuccess ||
        hipMemcpyAsync(g_noalloc_host, d->rows, bytes, hipMemcpyDeviceToHost, d->stream) !=
            hipSuccess ||
        hipStreamSynchronize(d->stream) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    float ms = 0.f;
    RefineState& s = state();
    {
        std::lock_guard<std::mutex> stats_lock(s.mu);
        if (hipEventElapsedTime(&ms, d->ev0, d->ev1) == hipSuccess) s.last_ns = (int64_t)(ms * 1e6);
        s.last_evals = (int64_t)count * a.noff * a.lv_n;
    }
    std::copy(g_noalloc_host, g_noalloc_host + count, rows);
    return 0;
}

void release_refine() {
    RefineState& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    if (s.d_rows) (void)hipFree(s.d_rows);
    if (s.d_seeds) (void)hipFree(s.d_seeds);
    if (s.ev0) (void)hipEventDestroy(s.ev0);
    if (s.ev1) (void)hipEventDestroy(s.ev1);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s.d_rows = nullptr;
    s.d_seeds = nullptr;
    s.cap = 0;
    s.ev0 = s.ev1 = nullptr;
    s.stream = nullptr;
}

}  // namespace slime
