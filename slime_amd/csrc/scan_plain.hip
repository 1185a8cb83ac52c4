// Plain cluster scan: every chunk gets its own LCG evaluation.
//
// One workgroup (4 wave64s) owns a tile of 1008 x tile_h centers. It walks the
// tile along z, four chunk rows per step:
//   1. each wave evaluates one 1024-chunk row (x = tile_x0-8 .. tile_x0+1015);
//      lane l takes chunk 64*j + l, and the wave's compare mask is row word j
//      (reference: ballot rows, SlimeCoreGPU.cu:412-452);
//   2. the rows land in a 24-row LDS ring (20 live rows + 4 being written);
//   3. thread g owns the 4x4 center group at columns 4g..4g+3 and keeps the
//      slime count of its 20x20 union square as a rolling sum of 4-row strips
//      (reference: group4 bound, SlimeCoreGPU.cu:812-918);
//   4. a group whose square reaches min_size gets the exact 221-cell count of
//      its 16 centers from the ring (SlimeCore.cpp:526-533), which feeds the
//      score histogram and the hit buffer.
// The ring holds first-draw bits (next(31) % 10 == 0), computed from Java's
// LCG directly (first_y of the 48-bit seed sum), with none of the banded
// kernels' phase/band arithmetic: an independent second path for the parity
// tests. The engine recounts the few centers near a chunk whose first draw
// Java rejects.
#include "scan.h"
#include "slime_math.h"

namespace slime {
namespace {

constexpr int kWords = 16;                  // 64-bit words per bitmap row
constexpr int kRowDwords = 2 * kWords + 2;  // + 2 pad dwords for window reads
constexpr int kRing = 24;
constexpr int kThreads = 256;
constexpr int kGroupCols = kTileOut / 4;    // 252

static_assert(kTileOut + 16 == kWords * 64, "tile row must fill the words");

// 20 bits of a ring row starting at bit `bit`.
__device__ __forceinline__ uint32_t window20(const uint32_t* row, int bit) {
    const int d = bit >> 5;
    return __builtin_amdgcn_alignbit(row[d + 1], row[d], bit & 31) & 0xFFFFFu;
}

// Exact disk counts of the 16 centers of a group from 20 ring rows starting at
// generated row `first`. cnt[4*j + i] is center (gx + i, gz + j); the window
// bit 8 + i is the center column.
__device__ __forceinline__ void exact_counts(const uint32_t (*ring)[kRowDwords], int first,
                                             int gbit, uint32_t cnt[16]) {
#pragma unroll
    for (int q = 0; q < 16; ++q) cnt[q] = 0;
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        const uint32_t w = window20(ring[(first + r) % kRing], gbit);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int dz = r - 8 - j;
            if (dz < -8 || dz > 8) continue;
            const int hw = disk_half_width(dz < 0 ? -dz : dz);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t mask = ((1u << (2 * hw + 1)) - 1u) << (8 + i - hw);
                cnt[4 * j + i] += __popc(w & mask);
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void scan_plain_kernel(ScanParams p) {
    __shared__ uint32_t ring[kRing][kRowDwords];
    __shared__ uint32_t lhist[kHistBins];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int i = tid; i < kHistBins; i += kThreads) lhist[i] = 0;
    if (tid < kRing) {
        ring[tid][2 * kWords] = 0;
        ring[tid][2 * kWords + 1] = 0;
    }

    const int tile = blockIdx.x;
    const int tx = tile % p.tiles_x;
    const int tz = tile / p.tiles_x;
    const int off_x = tx * kTileOut;  // first center column offset in the region
    const int off_z = tz * p.tile_h;
    // Coordinates wrap like Java ints; region checks use offsets.
    const uint32_t chunk_x0 = (uint32_t)p.x0 + (uint32_t)off_x - 8u;
    const uint32_t chunk_z0 = (uint32_t)p.z0 + (uint32_t)off_z - 8u;

    uint64_t xs[kWords];  // seed + xt(x) of the lane's chunk in each row word
#pragma unroll
    for (int j = 0; j < kWords; ++j)
        xs[j] = (uint64_t)p.seed + x_term((int32_t)(chunk_x0 + 64u * j + (uint32_t)lane));

    const int g = tid;
    const int gbit = 4 * g;
    uint32_t square = 0;   // 20x20 count of the group's current union square
    uint64_t strips = 0;   // 4-row strip counts of the last batches, a byte each
    const int nbatch = p.tile_h / 4 + 4;

    for (int k = 0; k < nbatch; ++k) {
        const int row = 4 * k + wave;
        const uint64_t Z = z_term((int32_t)(chunk_z0 + (uint32_t)row));
        uint32_t mine = 0;
#pragma unroll
        for (int j = 0; j < kWords; ++j) {
            const bool hit = first_draw_hit(first_y(xs[j] + Z));
            const uint64_t m = __ballot(hit);
            mine = lane == 2 * j ? (uint32_t)m : mine;
            mine = lane == 2 * j + 1 ? (uint32_t)(m >> 32) : mine;
        }
        if (lane < 2 * kWords) ring[row % kRing][lane] = mine;
        __syncthreads();

        if (g < kGroupCols) {
            uint32_t c4 = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) c4 += __popc(window20(ring[(4 * k + r) % kRing], gbit));
            const uint32_t leaving = (uint32_t)(strips >> 32) & 0xFFu;
            square = square + c4 - leaving;
            strips = (strips << 8) | c4;
            const int m = k - 4;
            const int gx_off = off_x + gbit;
            const int gz_off = off_z + 4 * m;
            if (m >= 0 && square >= (uint32_t)p.min_size && gx_off < p.w && gz_off < p.h &&
                4 * m < p.tile_h) {
                uint32_t cnt[16];
                exact_counts(ring, 4 * m, gbit, cnt);
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int i = q & 3, j = q >> 2;
                    const uint32_t c = cnt[q];
                    if (c < (uint32_t)p.min_size || c > (uint32_t)p.max_size) continue;
                    if (gx_off + i >= p.w || gz_off + j >= p.h) continue;
                    const int32_t cx = (int32_t)((uint32_t)p.x0 + (uint32_t)(gx_off + i));
                    const int32_t cz = (int32_t)((uint32_t)p.z0 + (uint32_t)(gz_off + j));
                    if (p.use_annulus) {
                        const int64_t dx = (int64_t)cx - p.acx, dz = (int64_t)cz - p.acz;
                        if (dx * dx + dz * dz < p.rd_min_sq) continue;
                    }
                    atomicAdd(&lhist[c], 1u);
                    if (c >= (uint32_t)p.emit_min) {
                        const unsigned long long idx = atomicAdd(p.emitted, 1ull);
                        if (idx < p.hit_cap) p.hits[idx] = Hit{(int32_t)c, cx, cz, 0};
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < kHistBins; i += kThreads)
        if (lhist[i]) atomicAdd(&p.hist[i], (unsigned long long)lhist[i]);
}

}  // namespace

hipError_t launch_scan_plain(const ScanParams& p, hipStream_t stream) {
    if (p.tiles_total <= 0) return hipSuccess;
    hipLaunchKernelGGL(scan_plain_kernel, dim3(p.tiles_total), dim3(kThreads), 0, stream, p);
    return hipGetLastError();
}

}  // namespace slime
