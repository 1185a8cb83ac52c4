// Persistent multi-search sweep: tiles of many banded jobs claimed from a
// device work queue by a fixed grid of workgroups.
//
// The queue is an ordered job list with one claim cursor per job. Wave 0 of a
// workgroup looks at 64 jobs at once from its current position, takes the
// first one whose request is active and whose cursor is below its tile count,
// and claims a tile with one atomic add; positions only move forward, so
// every workgroup exits once it has passed the end of the list (or the
// launch's claim budget is spent). There is one kernel per class (bands,
// x-phases); its tiles run the same code as the per-search kernels
// (band_tile.h).
#include <utility>

#include "band_tile.h"

namespace slime {
namespace {

using band::kThreads;

// Claims one tile for the workgroup (called by wave 0, uniform result).
__device__ bool claim_tile(const SweepArgs& a, int& pos, int& job, int& tile) {
    const int lane = threadIdx.x & 63;
    while (pos < a.norder) {
        const int idx = pos + lane;
        int j = -1;
        bool open = false;
        if (idx < a.norder) {
            j = a.order[idx];
            const SweepJob& jb = a.jobs[j];
            open = a.status[jb.req] == kReqActive &&
                   __hip_atomic_load(&a.cursor[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       (unsigned int)jb.p.tiles_total;
        }
        const uint64_t m = __ballot(open);
        if (m == 0) {
            pos += 64;
            continue;
        }
        const int first = __builtin_ctzll(m);
        pos += first;  // the jobs before it are finished or not active
        unsigned int t = 0;
        if (lane == first) t = atomicAdd(&a.cursor[j], 1u);
        t = __shfl(t, first);
        const int jj = __shfl(j, first);
        if (t < (unsigned int)a.jobs[jj].p.tiles_total) {
            job = jj;
            tile = (int)t;
            return true;
        }
        pos += 1;  // another workgroup took the last tile
    }
    return false;
}

// The persistent kernel of one class: a workgroup claims a tile, scans it,
// flushes its histogram and claims the next, until the queue or the budget is
// exhausted. One kernel per class keeps each class's register allocation
// (a runtime switch over all classes in one kernel spills).
// Waves per SIMD each class is built for: the occupancy its per-search kernel
// reaches with the same tile code (fewer bands need fewer registers).
// The union-bound kernels keep their class's occupancy where the extra
// rolling state fits. The 2-band, two-phase class runs at four waves in both
// forms (at five, with the claim loop around the tile and, in the union-bound
// form, the matrix-core window sums, it spilled). The three-phase kernels
// (union bound only) run at four (the 5-band one at three: at four it spilled
// 16 bytes), the 12- and 14-band union-bound kernels at two (their one-tile
// form needs up to 167 VGPRs before the claim loop).
constexpr int kSweepWaves[kSweepKernels] = {6, 5, 5, 5, 5, 4, 4, 4, 4, 3, 3, 6, 5, 4,
                                            4, 5, 4, 4, 4, 3, 3, 4, 4, 4, 3, 2, 3, 2};

template <int K>
__global__ __launch_bounds__(kThreads, kSweepWaves[K]) void sweep_kernel(SweepArgs a) {
    constexpr int C = sweep_class_of_kernel(K);
    constexpr int NB = kSweepClass[C].nb, PH = kSweepClass[C].ph;
    constexpr bool UB = sweep_kernel_union(K);
    __shared__ band::TileLds<NB, PH> lds;
    __shared__ int s_job, s_tile;
    for (int i = threadIdx.x; i < kHistBins; i += kThreads) lds.lhist[i] = 0;
    int pos = 0;
    for (;;) {
        if (threadIdx.x < 64) {
            unsigned int c = 0;
            if (threadIdx.x == 0) c = atomicAdd(a.claims, 1u);
            c = __shfl(c, 0);
            int job = -1, tile = 0;
            const bool ok = c < a.budget && claim_tile(a, pos, job, tile);
            if (threadIdx.x == 0) {
                s_job = ok ? job : -1;
                s_tile = tile;
            }
        }
        __syncthreads();
        const int job = __builtin_amdgcn_readfirstlane(s_job);
        const int tile = __builtin_amdgcn_readfirstlane(s_tile);
        if (job < 0) break;
        // read in place (scalar loads as needed) rather than copied into registers
        const BandParams& p = a.jobs[job].p;
        // a copy of the thread index the compiler cannot see through: the
        // tile's per-lane constants are then computed per tile instead of
        // being hoisted out of the claim loop (live across tiles, they cost
        // registers and waves)
        int tid;
        asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
        band::band_tile<NB, PH, UB>(p, p.emit_min, tile, lds, tid);
        band::flush_hist(lds.lhist, p.hist);
        if (threadIdx.x == 0) atomicAdd(&a.done[job], 1u);
        __syncthreads();
    }
}

// A launch of kernel K without claims: one workgroup per tile of the jobs
// a.order[0..n), job i's tiles start[i] + ... being the blocks
// [first[i], first[i+1]).
// Each workgroup finds its job by binary search and runs the tile as the
// per-search kernels do (same register budget: no claim loop around it).
template <int K>
__global__ __launch_bounds__(kThreads, 3) void batch_kernel(BatchArgs a) {
    constexpr int C = sweep_class_of_kernel(K);
    constexpr int NB = kSweepClass[C].nb, PH = kSweepClass[C].ph;
    constexpr bool UB = sweep_kernel_union(K);
    __shared__ band::TileLds<NB, PH> lds;
    for (int i = threadIdx.x; i < kHistBins; i += kThreads) lds.lhist[i] = 0;
    const int64_t blk = blockIdx.x;
    int lo = 0, hi = a.n - 1;  // the last i with first[i] <= blk
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.first[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const int job = a.order[lo];
    const BandParams p = a.jobs[job].p;
    band::band_tile<NB, PH, UB>(p, p.emit_min, a.start[lo] + (int)(blk - a.first[lo]), lds,
                                (int)threadIdx.x);
    band::flush_hist(lds.lhist, p.hist);
    if (threadIdx.x == 0) atomicAdd(&a.done[job], 1u);
}

template <int K>
hipError_t launch_batch_kernel(const BatchArgs& a, int64_t blocks, hipStream_t stream) {
    hipLaunchKernelGGL(batch_kernel<K>, dim3((unsigned)blocks), dim3(kThreads), 0, stream, a);
    return hipGetLastError();
}

template <int C>
hipError_t launch_class(const SweepArgs& a, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(sweep_kernel<C>, dim3(grid), dim3(kThreads), 0, stream, a);
    return hipGetLastError();
}

template <int C>
int blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sweep_kernel<C>, kThreads, 0) !=
            hipSuccess ||
        n < 1)
        return 1;
    return n;
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

int sweep_blocks_per_cu(int kernel) {
    switch (kernel) {
        case 0: return blocks_per_cu<0>();
        case 1: return blocks_per_cu<1>();
        case 2: return blocks_per_cu<2>();
        case 3: return blocks_per_cu<3>();
        case 4: return blocks_per_cu<4>();
        case 5: return blocks_per_cu<5>();
        case 6: return blocks_per_cu<6>();
        case 7: return blocks_per_cu<7>();
        case 8: return blocks_per_cu<8>();
        case 9: return blocks_per_cu<9>();
        case 10: return blocks_per_cu<10>();
        case 11: return blocks_per_cu<11>();
        case 12: return blocks_per_cu<12>();
        case 13: return blocks_per_cu<13>();
        case 14: return blocks_per_cu<14>();
        case 15: return blocks_per_cu<15>();
        case 16: return blocks_per_cu<16>();
        case 17: return blocks_per_cu<17>();
        case 18: return blocks_per_cu<18>();
        case 19: return blocks_per_cu<19>();
        case 20: return blocks_per_cu<20>();
        case 21: return blocks_per_cu<21>();
        case 22: return blocks_per_cu<22>();
        case 23: return blocks_per_cu<23>();
        case 24: return blocks_per_cu<24>();
        case 25: return blocks_per_cu<25>();
        case 26: return blocks_per_cu<26>();
        case 27: return blocks_per_cu<27>();
        default: return 1;
    }
}

template <int... K>
int batch_blocks_per_cu_of(int kernel, std::integer_sequence<int, K...>) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    (void)((kernel == K
                ? (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<K>, kThreads, 0),
                   true)
                : false) ||
           ...);
    return e == hipSuccess && n >= 1 ? n : 1;
}

int batch_blocks_per_cu(int kernel) {
    return batch_blocks_per_cu_of(kernel, std::make_integer_sequence<int, kSweepKernels>{});
}

hipError_t launch_sweep(const SweepArgs& a, int kernel, int grid, hipStream_t stream) {
    if (a.norder <= 0 || grid <= 0) return hipSuccess;
    switch (kernel) {
        case 0: return launch_class<0>(a, grid, stream);
        case 1: return launch_class<1>(a, grid, stream);
        case 2: return launch_class<2>(a, grid, stream);
        case 3: return launch_class<3>(a, grid, stream);
        case 4: return launch_class<4>(a, grid, stream);
        case 5: return launch_class<5>(a, grid, stream);
        case 6: return launch_class<6>(a, grid, stream);
        case 7: return launch_class<7>(a, grid, stream);
        case 8: return launch_class<8>(a, grid, stream);
        case 9: return launch_class<9>(a, grid, stream);
        case 10: return launch_class<10>(a, grid, stream);
        case 11: return launch_class<11>(a, grid, stream);
        case 12: return launch_class<12>(a, grid, stream);
        case 13: return launch_class<13>(a, grid, stream);
        case 14: return launch_class<14>(a, grid, stream);
        case 15: return launch_class<15>(a, grid, stream);
        case 16: return launch_class<16>(a, grid, stream);
        case 17: return launch_class<17>(a, grid, stream);
        case 18: return launch_class<18>(a, grid, stream);
        case 19: return launch_class<19>(a, grid, stream);
        case 20: return launch_class<20>(a, grid, stream);
        case 21: return launch_class<21>(a, grid, stream);
        case 22: return launch_class<22>(a, grid, stream);
        case 23: return launch_class<23>(a, grid, stream);
        case 24: return launch_class<24>(a, grid, stream);
        case 25: return launch_class<25>(a, grid, stream);
        case 26: return launch_class<26>(a, grid, stream);
        case 27: return launch_class<27>(a, grid, stream);
        default: return hipErrorInvalidValue;
    }
}
static_assert(kSweepKernels == 28, "one case per kernel");

hipError_t launch_batch(const BatchArgs& a, int kernel, int64_t blocks, hipStream_t stream) {
    if (a.n <= 0 || blocks <= 0) return hipSuccess;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    switch (kernel) {
        case 0: return launch_batch_kernel<0>(a, blocks, stream);
        case 1: return launch_batch_kernel<1>(a, blocks, stream);
        case 2: return launch_batch_kernel<2>(a, blocks, stream);
        case 3: return launch_batch_kernel<3>(a, blocks, stream);
        case 4: return launch_batch_kernel<4>(a, blocks, stream);
        case 5: return launch_batch_kernel<5>(a, blocks, stream);
        case 6: return launch_batch_kernel<6>(a, blocks, stream);
        case 7: return launch_batch_kernel<7>(a, blocks, stream);
        case 8: return launch_batch_kernel<8>(a, blocks, stream);
        case 9: return launch_batch_kernel<9>(a, blocks, stream);
        case 10: return launch_batch_kernel<10>(a, blocks, stream);
        case 11: return launch_batch_kernel<11>(a, blocks, stream);
        case 12: return launch_batch_kernel<12>(a, blocks, stream);
        case 13: return launch_batch_kernel<13>(a, blocks, stream);
        case 14: return launch_batch_kernel<14>(a, blocks, stream);
        case 15: return launch_batch_kernel<15>(a, blocks, stream);
        case 16: return launch_batch_kernel<16>(a, blocks, stream);
        case 17: return launch_batch_kernel<17>(a, blocks, stream);
        case 18: return launch_batch_kernel<18>(a, blocks, stream);
        case 19: return launch_batch_kernel<19>(a, blocks, stream);
        case 20: return launch_batch_kernel<20>(a, blocks, stream);
        case 21: return launch_batch_kernel<21>(a, blocks, stream);
        case 22: return launch_batch_kernel<22>(a, blocks, stream);
        case 23: return launch_batch_kernel<23>(a, blocks, stream);
        case 24: return launch_batch_kernel<24>(a, blocks, stream);
        case 25: return launch_batch_kernel<25>(a, blocks, stream);
        case 26: return launch_batch_kernel<26>(a, blocks, stream);
        case 27: return launch_batch_kernel<27>(a, blocks, stream);
        default: return hipErrorInvalidValue;
    }
}

// One atomic per wave: the lanes that qualify take consecutive slots from a
// wave-wide ballot.
__global__ __launch_bounds__(256) void select_hits_kernel(const Hit* __restrict__ hits, int64_t n,
                                                          int32_t min_size, Hit* __restrict__ out,
                                                          uint32_t cap, unsigned int* count) {
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        const bool take = i < n && hits[i].size >= min_size;
        const unsigned long long m = __ballot(take);
        if (m == 0) continue;
        unsigned int base = 0;
        if (lane == 0) base = atomicAdd(count, (unsigned int)__popcll(m));
        base = (unsigned int)__shfl((int)base, 0);
        if (take) {
            const unsigned int idx = base + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
            if (idx < cap) out[idx] = hits[i];
        }
    }
}

hipError_t launch_select_hits(const Hit* hits, int64_t n, int32_t min_size, Hit* out,
                              uint32_t cap, unsigned int* count, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(select_hits_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, hits, n,
                       min_size, out, cap, count);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

}  // namespace slime
