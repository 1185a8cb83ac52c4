// A/B of the banded scan's counting stage: the rolling 20x20 window count of
// every 4x4 group (minus the 32 corner cells the union of disks never holds),
// per band-phase, counted (0) on the VALU by plane words, as the scan kernels
// count, or (1) with the x-direction window sum as an int8 contraction on the
// matrix cores (v_mfma_i32_16x16x64_i8).
//
// Both methods read the same bitplane ring in LDS, laid out as the scan's
// (an E plane and one hit plane per band, 512 chunk columns a row, 24 rows),
// filled once from a hash so that the work measured is the counting alone,
// one strip of 4 rows per step, with a workgroup barrier per step as the
// scan has after each row batch. The vertical part (per-nibble popcounts of
// the strip's rows, 20-row column counts as bytes, the corner rows) is the
// same instruction sequence in both; they differ in how the 5-column window
// sums and the threshold are formed:
//   0: 16-bit lane sums of the byte counts with the next word's counts taken
//      from the next lane (DPP), packed threshold bits;
//   1: the counts and corner counts of 16 band-phases x 32 nibble columns as
//      the A operand (a lane's 16 bytes: 4 even and 4 odd column counts, 4
//      even and 4 odd corner counts of one plane word), a constant 0/+1/-1
//      window matrix as B, the threshold bias as C: one instruction gives 16
//      band-phases x 16 groups. A wave takes 32 nibble columns; the window of
//      its last 16 groups reaches into the next wave's columns, so the counts
//      also go through LDS (one 16-byte write and read per lane and step,
//      and a second barrier).
// Results: per workgroup a wrapping sum of a weight of (strip, band-phase,
// group) over the windows that reach min_size, and their number, so that both
// methods can be checked against a CPU count of the same hashed planes.
// No counterpart in the reference (its binary tensor-core popcounts are
// cuda_x2_z15/tensor_cold.py:4-39); measurement only.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <new>

#include "../../include/slimefinder.h"
#include "band_tile.h"

namespace {

using slime::band::from_next_lane16;
using slime::band::nibble_counts;

constexpr int kAbThreads = 256;
constexpr int kAbRing = 24;
constexpr int kAbMaxPlanes = 9;                  // E + up to 8 bands
constexpr int kAbRowDw = kAbMaxPlanes * 16 + 2;  // 512 bits a plane + pad
constexpr int kAbGroups = 124;                   // groups of a 496-center tile row
constexpr int kAbWords = 16;                     // plane words a row
constexpr int kAbXchWords = kAbWords + 2;        // + 2 zero words past the row

typedef int v4i __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint32_t ab_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ inline uint32_t ab_weight(uint32_t m, uint32_t q, uint32_t g) {
    return m * 0x9E3779B1u + q * 0x85EBCA77u + g * 0xC2B2AE3Du;
}

// Rolling vertical state of one (band-phase, plane word): 20-row counts of
// the word's 8 nibble columns as bytes (even columns, odd columns), the last
// five strips in the same form, and rows 0-1 of the last four strips.
struct Vert {
    uint32_t we, wo, he[5], ho[5], hq[4];
};

__device__ __forceinline__ void vert_init(Vert& v) {
    v.we = v.wo = 0;
#pragma unroll
    for (int j = 0; j < 5; ++j) v.he[j] = v.ho[j] = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) v.hq[j] = 0;
}

// Adds strip k (rows 4k..4k+3 of the stream; stream row s is ring row s % 24)
// and gives the corner counts of the window that ends with it: rows 0-1 (of
// the strip four steps back) plus rows 18-19 (this strip's rows 2-3), per
// nibble column, as bytes (pe even columns, po odd).
__device__ __forceinline__ void vert_step(Vert& v, const uint32_t* ring, int k, int wd, int plane,
                                          uint32_t flip, uint32_t& pe, uint32_t& po) {
    uint32_t nhead = 0, ntail = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t* rowp = ring + ((4 * k + r) % kAbRing) * kAbRowDw;
        const uint32_t c = nibble_counts(rowp[plane * 16 + wd] & (rowp[wd] ^ flip));
        if (r < 2) nhead += c;
        else ntail += c;
    }
    const uint32_t se = (nhead & 0x0F0F0F0Fu) + (ntail & 0x0F0F0F0Fu);
    const uint32_t so = ((nhead >> 4) & 0x0F0F0F0Fu) + ((ntail >> 4) & 0x0F0F0F0Fu);
    v.we += se - v.he[0];
    v.wo += so - v.ho[0];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v.he[j] = v.he[j + 1];
        v.ho[j] = v.ho[j + 1];
    }
    v.he[4] = se;
    v.ho[4] = so;
    const uint32_t old = v.hq[0];
#pragma unroll
    for (int j = 0; j < 3; ++j) v.hq[j] = v.hq[j + 1];
    v.hq[3] = nhead;
    pe = (ntail & 0x0F0F0F0Fu) + (old & 0x0F0F0F0Fu);
    po = ((ntail >> 4) & 0x0F0F0F0Fu) + ((old >> 4) & 0x0F0F0F0Fu);
}

// Fills the ring from the hash: plane 0 (E) one hash a dword, hit planes the
// AND of three (a bit is set with probability 1/8). Pads are zero.
template <int NB>
__device__ void fill_ring(uint32_t* ring, uint32_t seed) {
    for (int i = threadIdx.x; i < kAbRing * kAbRowDw; i += kAbThreads) {
        const int r = i / kAbRowDw, o = i % kAbRowDw, p = o / 16, d = o % 16;
        uint32_t v = 0;
        if (p <= NB) {
            const uint32_t idx =
                ((((uint32_t)blockIdx.x * kAbRing + (uint32_t)r) * kAbMaxPlanes + (uint32_t)p) * 16u +
                 (uint32_t)d) * 3u + seed;
            v = ab_hash(idx);
            if (p > 0) v &= ab_hash(idx + 1u) & ab_hash(idx + 2u);
        }
        ring[i] = v;
    }
}

// Method 0: VALU word counting with the scan kernels' thread roles (16
// threads per band-phase, one plane word each; waves past the last band
// idle in counting).
template <int NB>
__global__ __launch_bounds__(kAbThreads, 3) void window_valu_kernel(int32_t steps, int32_t min_size,
                                                                  uint32_t seed, uint32_t* out) {
    __shared__ uint32_t ring[kAbRing * kAbRowDw];
    fill_ring<NB>(ring, seed);
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int wd = tid & 15, wb = tid >> 5, wph = (tid >> 4) & 1;
    constexpr int kWThreads = 32 * NB;
    const bool active = wb < NB;
    const int wbc = active ? wb : 0;
    const uint32_t flip = ((wbc ^ wph) & 1) ? 0u : 0xFFFFFFFFu;
    const uint32_t q = (uint32_t)(2 * wbc + wph);
    const uint32_t valid = active ? (wd == kAbWords - 1 ? 0x0Fu : 0xFFu) : 0u;
    const uint32_t wbias = (uint32_t)(0x8000 - min_size) * 0x10001u;
    Vert v;
    vert_init(v);
    uint32_t cs = 0, np = 0;
    __syncthreads();
    for (int k = 0; k < steps; ++k) {
        __syncthreads();  // the scan's barrier after each row batch
        if (wave * 64 < kWThreads) {
            uint32_t pe, po;
            vert_step(v, ring, k, wd, wbc + 1, flip, pe, po);
            const uint32_t ne = from_next_lane16(v.we), no = from_next_lane16(v.wo);
            const uint32_t w0 = v.we & 0x00FF00FFu, w2 = (v.we >> 8) & 0x00FF00FFu;
            const uint32_t w1 = v.wo & 0x00FF00FFu, w3 = (v.wo >> 8) & 0x00FF00FFu;
            const uint32_t n0 = ne & 0x00FF00FFu, n2 = (ne >> 8) & 0x00FF00FFu;
            const uint32_t n1 = no & 0x00FF00FFu, n3 = (no >> 8) & 0x00FF00FFu;
            uint32_t q0 = w0 + w1 + w2 + w3 + __builtin_amdgcn_alignbit(n0, w0, 16);
            uint32_t q1 = q0 - w0 + __builtin_amdgcn_alignbit(n1, w1, 16);
            uint32_t q2 = q1 - w1 + __builtin_amdgcn_alignbit(n2, w2, 16);
            uint32_t q3 = q2 - w2 + __builtin_amdgcn_alignbit(n3, w3, 16);
            const uint32_t npe = from_next_lane16(pe), npo = from_next_lane16(po);
            const uint32_t p0 = pe & 0x00FF00FFu, p2 = (pe >> 8) & 0x00FF00FFu;
            const uint32_t p1 = po & 0x00FF00FFu, p3 = (po >> 8) & 0x00FF00FFu;
            q0 -= p0 + __builtin_amdgcn_alignbit(npe & 0x00FF00FFu, p0, 16);
            q1 -= p1 + __builtin_amdgcn_alignbit(npo & 0x00FF00FFu, p1, 16);
            q2 -= p2 + __builtin_amdgcn_alignbit((npe >> 8) & 0x00FF00FFu, p2, 16);
            q3 -= p3 + __builtin_amdgcn_alignbit((npo >> 8) & 0x00FF00FFu, p3, 16);
            const uint32_t c0 = q0 + wbias, c1 = q1 + wbias, c2 = q2 + wbias, c3 = q3 + wbias;
            uint32_t pm = ((c0 >> 15) & 1u) | ((c1 >> 14) & 2u) | ((c2 >> 13) & 4u) |
                          ((c3 >> 12) & 8u) | ((c0 >> 27) & 16u) | ((c1 >> 26) & 32u) |
                          ((c2 >> 25) & 64u) | ((c3 >> 24) & 128u);
            pm &= valid;
            if (k < 4) pm = 0;
            while (pm) {  // rare at scan thresholds
                const int i = __builtin_ctz(pm);
                pm &= pm - 1u;
                cs += ab_weight((uint32_t)(k - 4), q, (uint32_t)(8 * wd + i));
                ++np;
            }
        }
    }
    atomicAdd(&out[2 * blockIdx.x], cs);
    atomicAdd(&out[2 * blockIdx.x + 1], np);
}

// Method 1: int8 MFMA window sums. 2 NB band-phases (QN) x H column ranges
// make the 16 rows of A (H = 16 / QN); lane l of wave w holds row m = l % 16
// (band-phase m % QN, range m / QN) over plane word
// (m / QN) * (16 / H) + 4 w + l / 16. Waves past 4 / H idle in counting.
template <int NB>
__global__ __launch_bounds__(kAbThreads, 3) void window_mfma_kernel(int32_t steps, int32_t min_size,
                                                                  uint32_t seed, uint32_t* out) {
    constexpr int QN = 2 * NB;
    constexpr int H = 16 / QN;
    constexpr int kRangeWords = kAbWords / H;
    constexpr int kWaves = 4 / H;
    static_assert(QN * H == 16 && kWaves * H == 4, "2, 4 or 8 bands");
    __shared__ uint32_t ring[kAbRing * kAbRowDw];
    __shared__ v4i xch[kAbXchWords * 16];
    fill_ring<NB>(ring, seed);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, j = lane >> 4;
    const int qm = m % QN;
    const int b = qm >> 1, ph = qm & 1;
    const int wd = (m / QN) * kRangeWords + 4 * wave + j;  // < 16 for counting waves
    const uint32_t flip = ((b ^ ph) & 1) ? 0u : 0xFFFFFFFFu;
    for (int i = tid; i < 2 * 16; i += kAbThreads) xch[kAbWords * 16 + i] = v4i{0, 0, 0, 0};

    // B: lane l gives column n = l % 16 (group n of the wave's first group),
    // k-block l / 16 (plane word 4 w + l / 16 relative to the wave's first):
    // bytes 0-3 even column counts, 4-7 odd ones, 8-11 / 12-15 the corner
    // counts of the same columns. Window of group n: columns n..n+4, less the
    // corner counts of columns n and n+4.
    v4i bw;
    {
        const int n = lane & 15;
        uint32_t dw[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = 8 * j + 2 * e + (d & 1);
                int wgt;
                if (d < 2) wgt = (c >= n && c <= n + 4) ? 1 : 0;
                else wgt = (c == n || c == n + 4) ? -1 : 0;
                word |= (uint32_t)(uint8_t)(int8_t)wgt << (8 * e);
            }
            dw[d] = word;
        }
        bw = v4i{(int)dw[0], (int)dw[1], (int)dw[2], (int)dw[3]};
    }
    // C: 0x8000 - min_size, so that bit 15 of a sum says it reaches min_size
    const int bias = 0x8000 - min_size;
    const v4i cb = v4i{bias, bias, bias, bias};
    // D: lane l holds rows 4 (l / 16) + i (i = 0..3) of column l % 16: bit
    // 4 f + i of the pass mask is that row's group 16 f + l % 16 (f: 0 own
    // words, 1 two words on).
    uint32_t valid = 0;
    uint32_t dq[4], dg[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int md = 4 * j + i;
        dq[i] = (uint32_t)(md % QN);
        dg[i] = (uint32_t)(8 * ((md / QN) * kRangeWords + 4 * wave) + (lane & 15));
#pragma unroll
        for (int f = 0; f < 2; ++f)
            if ((int)dg[i] + 16 * f < kAbGroups) valid |= 1u << (4 * f + i);
    }
    if (wave >= kWaves) valid = 0;
    Vert v;
    vert_init(v);
    uint32_t cs = 0, np = 0;
    __syncthreads();
    for (int k = 0; k < steps; ++k) {
        __syncthreads();  // the scan's barrier after each row batch
        v4i a0 = v4i{0, 0, 0, 0};
        if (wave < kWaves) {
            uint32_t pe, po;
            vert_step(v, ring, k, wd, b + 1, flip, pe, po);
            a0 = v4i{(int)v.we, (int)v.wo, (int)pe, (int)po};
            xch[wd * 16 + qm] = a0;
        }
        __syncthreads();
        if (wave < kWaves) {
            // indexed by band-phase: with several column ranges, the rows
            // of another range hold the word
            const v4i a1 = xch[(wd + 2) * 16 + qm];
            const v4i d0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, bw, cb, 0, 0, 0);
            const v4i d1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, bw, cb, 0, 0, 0);
            const uint32_t lo = (((uint32_t)d0[0] & 0x8000u) | (((uint32_t)d0[1] & 0x8000u) << 1) |
                                 (((uint32_t)d0[2] & 0x8000u) << 2) |
                                 (((uint32_t)d0[3] & 0x8000u) << 3)) >> 15;
            const uint32_t hi = (((uint32_t)d1[0] & 0x8000u) | (((uint32_t)d1[1] & 0x8000u) << 1) |
                                 (((uint32_t)d1[2] & 0x8000u) << 2) |
                                 (((uint32_t)d1[3] & 0x8000u) << 3)) >> 11;
            uint32_t pm = (lo | hi) & valid;
            if (k < 4) pm = 0;
            while (pm) {  // rare at scan thresholds
                const int i = __builtin_ctz(pm);
                pm &= pm - 1u;
                cs += ab_weight((uint32_t)(k - 4), dq[i & 3], dg[i & 3] + 16u * (uint32_t)(i >> 2));
                ++np;
            }
        }
    }
    atomicAdd(&out[2 * blockIdx.x], cs);
    atomicAdd(&out[2 * blockIdx.x + 1], np);
}

template <int NB>
hipError_t launch_ab(int method, int blocks, int steps, int min_size, uint32_t seed, uint32_t* out,
                     hipStream_t s) {
    if (method == 0)
        hipLaunchKernelGGL(window_valu_kernel<NB>, dim3(blocks), dim3(kAbThreads), 0, s, steps,
                           min_size, seed, out);
    else
        hipLaunchKernelGGL(window_mfma_kernel<NB>, dim3(blocks), dim3(kAbThreads), 0, s, steps,
                           min_size, seed, out);
    return hipGetLastError();
}

hipError_t launch_ab_any(int method, int bands, int blocks, int steps, int min_size, uint32_t seed,
                         uint32_t* out, hipStream_t s) {
    switch (bands) {
        case 2: return launch_ab<2>(method, blocks, steps, min_size, seed, out, s);
        case 4: return launch_ab<4>(method, blocks, steps, min_size, seed, out, s);
        default: return launch_ab<8>(method, blocks, steps, min_size, seed, out, s);
    }
}

}  // namespace

extern "C" int32_t slime_window_ab(int32_t method, int32_t bands, int32_t blocks, int32_t steps,
                                   int32_t min_size, uint32_t seed, uint32_t* checksum,
                                   uint64_t* passes, int64_t* kernel_ns) {
    if ((method != 0 && method != 1) || (bands != 2 && bands != 4 && bands != 8) || blocks < 1 ||
        blocks > 65536 || steps < 5 || steps > (1 << 20) || min_size < 1 || min_size > 400)
        return -1;
    uint32_t* out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    const size_t bytes = sizeof(uint32_t) * 2 * (size_t)blocks;
    bool ok = hipMalloc(&out, bytes) == hipSuccess &&
              hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              // warm-up (code load), then the timed launch on cleared counters
              launch_ab_any(method, bands, blocks, 5, min_size, seed, out, s) == hipSuccess &&
              hipMemsetAsync(out, 0, bytes, s) == hipSuccess && hipEventRecord(e0, s) == hipSuccess &&
              launch_ab_any(method, bands, blocks, steps, min_size, seed, out, s) == hipSuccess &&
              hipEventRecord(e1, s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess &&
              hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    uint32_t cs = 0;
    uint64_t np = 0;
    if (ok) {
        uint32_t* h = new (std::nothrow) uint32_t[2 * (size_t)blocks];
        ok = h && hipMemcpy(h, out, bytes, hipMemcpyDeviceToHost) == hipSuccess;
        if (ok)
            for (int i = 0; i < blocks; ++i) {
                cs += h[2 * i];
                np += h[2 * i + 1];
            }
        delete[] h;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    if (!ok) {
        (void)hipGetLastError();
        return -1;
    }
    if (checksum) *checksum = cs;
    if (passes) *passes = np;
    if (kernel_ns) *kernel_ns = (int64_t)((double)ms * 1e6);
    return 0;
}
