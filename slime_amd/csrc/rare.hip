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

#include "scan.h"
#include "slime_math.h"

namespace slime {

namespace {

// Inverse of the Java LCG multiplier mod 2^64 (Newton iteration, odd input).
constexpr uint64_t inverse_mod_2_64(uint64_t a) {
    uint64_t x = a;  // correct to 3 bits for odd a
    for (int i = 0; i < 6; ++i) x *= 2 - a * x;
    return x;
}
constexpr uint64_t kJavaMulInv = inverse_mod_2_64(kJavaMul);
static_assert(((kJavaMul * kJavaMulInv) & kMask48) == 1, "LCG inverse");

std::vector<uint64_t> build_table() {
    std::vector<uint64_t> t;
    const uint64_t first = ((uint64_t)(0x7FFFFFFFu - 7u)) << 17;  // next(31) >= 2^31-8
    for (uint64_t st1 = first; st1 <= kMask48; ++st1) {
        uint32_t bits = (uint32_t)(st1 >> 17);
        bool first_draw = bits % 10u == 0u;
        uint64_t st = st1;
        uint32_t val = bits % 10u;
        while (bits - val + 9u > 0x7FFFFFFFu) {
            st = (st * kJavaMul + 11u) & kMask48;
            bits = (uint32_t)(st >> 17);
            val = bits % 10u;
        }
        if (first_draw == (val == 0u)) continue;
        const uint64_t st0 = ((st1 - 11u) * kJavaMulInv) & kMask48;
        t.push_back((st0 ^ kScramble) & kMask48);
    }
    std::sort(t.begin(), t.end());
    return t;
}

// The x (mod 2^32) with x*x*4987142 + x*5947611 == v (mod 2^32).
__device__ __forceinline__ uint32_t invert_x_low32(uint32_t v) {
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        const uint32_t f = x * x * 4987142u + x * 5947611u;
        x |= ((f ^ v) >> k & 1u) << k;
    }
    return x;
}

__device__ __forceinline__ int32_t lower_bound(const uint64_t* t, int32_t n, uint64_t key) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (t[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void find_mismatches_kernel(
    int64_t seed, int32_t x_first, int64_t cols, int32_t z_first, int64_t rows,
    const uint64_t* table, int32_t n, MismatchChunk* out, uint32_t cap, unsigned int* count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += stride) {
        const int32_t z = (int32_t)((uint32_t)z_first + (uint32_t)i);
        const uint64_t base = ((uint64_t)seed + z_term(z)) & kMask48;
        const uint64_t lo = (base - (1ull << 32)) & kMask48;  // xt >= -2^32
        int32_t k = lower_bound(table, n, lo);
        for (int32_t c = 0; c < n; ++c, ++k) {
            if (k >= n) k -= n;  // the window may wrap past 2^48
            const uint64_t off = (table[k] - lo) & kMask48;
            if (off >= (1ull << 33)) break;
            const int64_t d = (int64_t)off - (1ll << 32);  // required xt(x)
            const int32_t x = (int32_t)invert_x_low32((uint32_t)(uint64_t)d);
            const int64_t dx = (int64_t)x - (int64_t)x_first;
            if (dx < 0 || dx >= cols) continue;
            if ((int64_t)x_term(x) != d) continue;
            const unsigned int slot = atomicAdd(count, 1u);
            if (slot < cap) out[slot] = MismatchChunk{x, z};
        }
    }
}

__global__ __launch_bounds__(256) void recount_kernel(int64_t seed, CenterFix* fixes, int32_t n) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
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
    static const std::vector<uint64_t> table = build_table();
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

hipError_t launch_recount(int64_t seed, CenterFix* fixes, int32_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(recount_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, seed, fixes, n);
    return hipGetLastError();
}

}  // namespace slime
