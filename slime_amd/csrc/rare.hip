// First-draw mismatches: the chunks where "next(31) % 10 == 0" (what the
// banded scan tests) disagrees with Java's nextInt(10) == 0.
//
// Java redraws only when next(31) >= 2^31 - 8, i.e. when the first LCG output
// lies in the top 2^20 states of 2^48 (SlimeCore.cpp:54-56). Following each of
// those states through the redraw loop gives the 209,712 first-LCG states whose
// verdict flips; mapped back through the LCG step and the seed scramble they
// are a fixed, sorted table of seed sums s = seed + xt(x) + zt(z) (mod 2^48).
// The reference builds the analogous rejected-then-hit table with Python
// (cuda_x2_z15/generate_slime_rare_tables.py:15-33) and searches it per row
// (xp_support.py:153-214); here:
//   * per chunk row z, xt(x) lies in [-2^32, 2^32), so only table entries in a
//     2^33-wide window around seed + zt(z) can occur (about 6 per row);
//   * the low 32 bits of xt(x) are x*x*4987142 + x*5947611 (mod 2^32), a
//     permutation of Z/2^32 (its derivative is odd), so each entry names one
//     int32 x, found by lifting one bit at a time, and verified in 48 bits.
#include <algorithm>

#include "mismatch_table.h"
#include "scan.h"
#include "slime_math.h"

namespace slime {

namespace {

__global__ __launch_bounds__(256) void find_mismatches_kernel(
    int64_t seed, int32_t x_first, int64_t cols, int32_t z_first, int64_t rows,
    const uint64_t* table, int32_t n, MismatchChunk* out, uint32_t cap, unsigned int* count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += stride) {
        const int32_t z = (int32_t)((uint32_t)z_first + (uint32_t)i);
        int32_t xs[8];
        const int found = row_table_hits(seed, z, table, n, x_first, cols, xs, 8);
        if (found == 0) continue;
        if (found > 8) {  // never seen (~6 candidates per row, far fewer in range)
            atomicAdd(count, (unsigned int)(cap + 1));
            continue;
        }
        for (int f = 0; f < found; ++f) {
            const unsigned int slot = atomicAdd(count, 1u);
            if (slot < cap) out[slot] = MismatchChunk{xs[f], z};
        }
    }
}

__global__ __launch_bounds__(256) void find_mismatches_multi_kernel(
    const MismatchQuery* q, int32_t nq, int64_t total_rows, const uint64_t* table, int32_t n,
    TaggedMismatch* out, uint32_t cap, unsigned int* count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total_rows; i += stride) {
        int32_t lo = 0, hi = nq - 1;  // last query with row_begin <= i
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (q[mid].row_begin <= i) lo = mid; else hi = mid - 1;
        }
        const MismatchQuery& qq = q[lo];
        const int64_t r = i - qq.row_begin;
        if (r < 0 || r >= qq.rows) continue;
        const int32_t z = (int32_t)((uint32_t)qq.z_first + (uint32_t)r);
        int32_t xs[8];
        const int found = row_table_hits(qq.seed, z, table, n, qq.x_first, qq.cols, xs, 8);
        if (found == 0) continue;
        if (found > 8) {  // never seen (~6 candidates per row, far fewer in range)
            atomicAdd(count, (unsigned int)(cap + 1));
            continue;
        }
        for (int f = 0; f < found; ++f) {
            const unsigned int slot = atomicAdd(count, 1u);
            if (slot < cap) out[slot] = TaggedMismatch{xs[f], z, qq.tag};
        }
    }
}

__global__ __launch_bounds__(256) void recount_kernel(CenterFix* fixes, int32_t n) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t seed = fixes[i].seed;
    const int32_t cx = fixes[i].cx, cz = fixes[i].cz;
    int32_t fd = 0, exact = 0;
    for (int dz = -8; dz <= 8; ++dz) {
        const int hw = disk_half_width(dz < 0 ? -dz : dz);
        const uint64_t zt = z_term((int32_t)((uint32_t)cz + (uint32_t)dz));
        for (int dx = -hw; dx <= hw; ++dx) {
            const uint64_t s =
                (uint64_t)seed + x_term((int32_t)((uint32_t)cx + (uint32_t)dx)) + zt;
            const uint32_t y = first_y(s);
            const bool f = first_draw_hit(y);
            fd += f;
            exact += (y >= kRejectY) ? redraw_hit(s) : f;
        }
    }
    fixes[i].first_draw = fd;
    fixes[i].exact = exact;
}

}  // namespace

const std::vector<uint64_t>& mismatch_table() {
    static const std::vector<uint64_t> table = build_mismatch_table();
    return table;
}

hipError_t launch_find_mismatches(int64_t seed, int32_t x_first, int64_t cols, int32_t z_first,
                                  int64_t rows, const uint64_t* table, int32_t table_n,
                                  MismatchChunk* out, uint32_t cap, unsigned int* count,
                                  hipStream_t stream) {
    if (rows <= 0 || cols <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((rows + 255) / 256, 65536);
    hipLaunchKernelGGL(find_mismatches_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, seed,
                       x_first, cols, z_first, rows, table, table_n, out, cap, count);
    return hipGetLastError();
}

hipError_t launch_find_mismatches_multi(const MismatchQuery* q, int32_t nq, int64_t total_rows,
                                        const uint64_t* table, int32_t table_n,
                                        TaggedMismatch* out, uint32_t cap, unsigned int* count,
                                        hipStream_t stream) {
    if (nq <= 0 || total_rows <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((total_rows + 255) / 256, 65536);
    hipLaunchKernelGGL(find_mismatches_multi_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       q, nq, total_rows, table, table_n, out, cap, count);
    return hipGetLastError();
}

hipError_t launch_recount(CenterFix* fixes, int32_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(recount_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, fixes, n);
    return hipGetLastError();
}

}  // namespace slime
