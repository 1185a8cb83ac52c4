// One tile of the banded first-draw scan, as device code shared by the
// per-search kernels (one tile per workgroup) and the multi-search sweep
// kernel (tiles claimed from a work queue).
//
// Geometry (P = 2^17). With PH = 2 the region's centers are cut into x-blocks
// of 2P columns: block k pairs phase 0 = [X0 + 2kP, +P) with phase 1 = +P.
// With PH = 1 a single phase spans the region (no pairing, for regions
// narrower than P). Along z a tile covers NB bands z0 + b*P + s for slot rows
// s in [s_begin, s_end). A workgroup owns one x-tile (496 slot columns, 512
// chunk columns with the halo) and tile_h slot rows, and walks them four rows
// per step. Every LDS ring row holds 1 + NB bitplanes of 512 bits:
//   plane 0      E: lane parity e of the slot (which phase is even per band)
//   plane 1 + b  H_b: the first-draw hit of band b for the phase selected by e
// Band b, phase p is the bitmap H_b & (E if (b ^ p) odd else ~E).
//
// Arithmetic per slot (x = phase-0 column, s = slot row). With the seed sum
// s_{p,b} = X_p + Z_b and state0 = s ^ 0x5E434E432, split as 16 low bits u>>16
// and 32 high bits a; Y = (state1 >> 16) mod 2^32 = a*M + 5u + w, where
// u = (low16 << 16) and w = (u * 0xDEECE66D + (11 << 16)) >> 32 are common to
// all PH x NB chunks of the slot (their sums agree mod 2^16, and the carry into
// the high half too). Y bit 1 (the parity of next(31)) flips with p and with
// b, and Y bit 0 is common, so at most one phase per band can be even and
//   next(31) % 10 == 0  <=>  (Y - (Y & 1)) * 5^-1 <= (2^32 - 1) / 5
//                       <=>  a * (M * 5^-1) + [u + (w - (Y & 1)) * 5^-1] <= 858993459.
// The bracket is the per-slot bias; per band it is add, xor, multiply-add and
// a borrow shifted into the band's bitplane register. With PH = 1 the tested
// chunk is always phase 0 and the parity plane says whether it is even.
// (reference: x_phase.py:185-252 and benchmarks/algorithm-optimizations.en.md
// describe the same 2^17 parity relation for 2 X phases and up to 15 Z bands.)
//
// Counting follows the plain scan: per band-phase a rolling 20x20 square of
// each 4x4 group, a 332-cell union-of-disks bound, then the exact 221-cell
// count. Java's rejected first draws are not handled here; the engine
// recounts the few affected centers.
#pragma once

#include <type_traits>

#include "scan.h"
#include "slime_math.h"

namespace slime {
namespace band {

constexpr int kChunksPerLane = 8;                  // consecutive chunks a lane evaluates
constexpr int kRowBits = 64 * kChunksPerLane;      // 512 chunk columns per tile row
constexpr int kPlaneDw = kRowBits / 32;            // 16 dwords per bitplane row
constexpr int kRing = 24;                          // 20 live rows + 4 being written
constexpr int kThreads = 256;
constexpr int kGroupCols = kBandTileOut / 4;       // 124 groups per tile row

static_assert(kBandTileOut + 16 == kRowBits, "tile row = centers + halo");

__device__ __forceinline__ uint32_t window20(const uint32_t* row, int bit) {
    const int d = bit >> 5;
    return __builtin_amdgcn_alignbit(row[d + 1], row[d], bit & 31) & 0xFFFFFu;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [I0, N).
template <int I0, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I0 < N) {
        f(std::integral_constant<int, I0>{});
        static_for<I0 + 1, N>(f);
    }
}

// acc = 2 * acc + [v > limit]: the borrow of (limit - v) is shifted in through
// VCC by an add-with-carry. The carry hand-off between two VALU instructions
// is interlocked (checked on gfx950); the asm only pins the two-instruction form.
__device__ __forceinline__ void shift_in_above(uint32_t& acc, uint32_t v, uint32_t limit) {
    uint32_t scratch;
    asm("v_sub_co_u32 %1, vcc, %3, %2\n\t"
        "v_addc_co_u32 %0, vcc, %0, %0, vcc"
        : "+v"(acc), "=&v"(scratch)
        : "v"(v), "v"(limit)
        : "vcc");
}

template <int NB>
struct Ring {
    static constexpr int kPlanes = NB + 1;
    static constexpr int kRowDw = kPlanes * kPlaneDw + 2;  // + pad for window reads
    static constexpr int kDwords = kRing * kRowDw;
};

// Counting roles: two threads per 4x4 group, each with half of the bands (both
// phases), except one band: with two phases each thread takes one phase, with
// one phase a single thread counts.
template <int NB, int PH>
struct Roles {
    static constexpr bool kSplitPhase = NB == 1 && PH == 2;
    static constexpr int kPerGroup = (NB == 1 && PH == 1) ? 1 : 2;
    static constexpr int kBands = kSplitPhase ? 1 : (NB + kPerGroup - 1) / kPerGroup;
    static constexpr int kPhases = kSplitPhase ? 1 : PH;
    static constexpr int kOwned = kBands * kPhases;
};

// Bit mask of the union of the 16 disks of a 4x4 group in group row r
// (0..19): centers sit at window bits 8..11, so the row spans bits
// [8 - h, 11 + h] with h the widest disk row any of the 4 center rows reaches.
__device__ __forceinline__ uint32_t union_row_mask(int r) {
    int h = -1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int dz = r - 8 - j;
        if (dz >= -8 && dz <= 8) h = max(h, disk_half_width(dz < 0 ? -dz : dz));
    }
    return ((1u << (4 + 2 * h)) - 1u) << (8 - h);
}

// The group of band b, phase ph whose 20x20 square passed: a second bound on
// the union of its 16 disks (332 cells), then the exact 221-cell count of each
// center, the histogram and the hit buffer.
template <int NB>
__device__ __forceinline__ void refine_group(const BandParams& p, int32_t emit_min,
                                             const uint32_t* ring, int first, int gbit, int b,
                                             int ph, int64_t ox0, int col, int32_t s_g,
                                             uint32_t* lhist) {
    using R = Ring<NB>;
    const uint32_t flip = ((b ^ ph) & 1) ? 0u : 0xFFFFFFFFu;  // phase mask is E or ~E
    auto window = [&](int r) {
        const uint32_t* row = ring + ((first + r) % kRing) * R::kRowDw;
        return window20(row + (b + 1) * kPlaneDw, gbit) & (window20(row, gbit) ^ flip);
    };
    uint32_t bound = 0;
#pragma unroll
    for (int r = 0; r < 20; ++r) bound += __popc(window(r) & union_row_mask(r));
    if (bound < (uint32_t)p.min_size) return;
    uint32_t cnt[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) cnt[q] = 0;
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        const uint32_t wr = window(r);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int dz = r - 8 - j;
            if (dz < -8 || dz > 8) continue;
            const int hw = disk_half_width(dz < 0 ? -dz : dz);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                cnt[4 * j + i] += __popc(wr & (((1u << (2 * hw + 1)) - 1u) << (8 + i - hw)));
        }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = q & 3, jr = q >> 2;
        const uint32_t c = cnt[q];
        if (c < (uint32_t)p.min_size || c > (uint32_t)p.max_size) continue;
        if (col + i >= p.phase_cols || ox0 + i >= p.w || ox0 + i < p.ox_min) continue;
        const int32_t s_c = s_g + jr;
        if (s_c >= p.s_end || (int64_t)b * kPeriod + s_c >= p.h_valid) continue;
        const int32_t cx = (int32_t)((uint32_t)p.x0 + (uint32_t)(ox0 + i));
        const int32_t cz =
            (int32_t)((uint32_t)p.z0 + (uint32_t)b * (uint32_t)kPeriod + (uint32_t)s_c);
        if (p.use_annulus) {
            const int64_t dx = (int64_t)cx - p.acx, dz = (int64_t)cz - p.acz;
            if (dx * dx + dz * dz < p.rd_min_sq) continue;
        }
        atomicAdd(&lhist[c], 1u);
        if (c >= (uint32_t)emit_min) {
            const unsigned long long idx = atomicAdd(p.emitted, 1ull);
            if (idx < p.hit_cap) p.hits[idx] = Hit{(int32_t)c, cx, cz, p.tag};
        }
    }
}

// Scans tile `tile` of the job into the caller's LDS ring (Ring<NB>::kDwords)
// and LDS histogram (kHistBins, accumulated). The caller synchronises the
// workgroup before the call when the ring was in use.
template <int NB, int PH>
__device__ __forceinline__ void band_tile(const BandParams& p, int32_t emit_min, int tile,
                                          uint32_t* ring, uint32_t* lhist) {
    static_assert(PH == 1 || PH == 2, "one or two x-phases");
    using R = Ring<NB>;
    using C = Roles<NB, PH>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid < kRing) {
        ring[tid * R::kRowDw + R::kPlanes * kPlaneDw] = 0;
        ring[tid * R::kRowDw + R::kPlanes * kPlaneDw + 1] = 0;
    }

    const int tx = tile % p.tiles_x;
    const int ts = tile / p.tiles_x;
    const int xblock = tx / p.tiles_per_block;
    const int xtile = tx % p.tiles_per_block;
    const int64_t block_off = (int64_t)xblock * 2 * kPeriod;  // center offset of phase 0
    const int32_t col0 = p.slot_lo + xtile * kBandTileOut;      // slot column offset in phase
    const uint32_t chunk_x0 = (uint32_t)p.x0 + (uint32_t)(block_off + col0) - 8u;

    // Lane l evaluates the chunks kChunksPerLane*l + j of every row.
    uint32_t xlo[kChunksPerLane], xh0[kChunksPerLane], xh1[PH == 2 ? kChunksPerLane : 1];
#pragma unroll
    for (int j = 0; j < kChunksPerLane; ++j) {
        const uint32_t xc = chunk_x0 + (uint32_t)(kChunksPerLane * lane + j);
        const uint64_t X0 = (uint64_t)p.seed + x_term((int32_t)xc);
        xlo[j] = (uint32_t)X0 << 16;
        xh0[j] = (uint32_t)(X0 >> 16);
        if constexpr (PH == 2) {
            const uint64_t X1 = (uint64_t)p.seed + x_term((int32_t)(xc + (uint32_t)kPeriod));
            xh1[j] = (uint32_t)(X1 >> 16);
        }
    }

    // Counting role: group g, bands [b_lo, b_lo + kBands), phases [ph_lo, +kPhases).
    const int g = C::kPerGroup == 2 ? tid >> 1 : tid;
    const int half = C::kPerGroup == 2 ? (tid & 1) : 0;
    const int b_lo = C::kSplitPhase ? 0 : half * C::kBands;
    const int ph_lo = C::kSplitPhase ? half : 0;
    const int gbit = 4 * g;
    const bool counts = g < kGroupCols;
    // Rolling state per owned band-phase q = kPhases*(b - b_lo) + (ph - ph_lo):
    // sq = square | (strip leaving next) << 16; hist = last four strips, a byte each.
    uint32_t sq[C::kOwned], hist[C::kOwned];
#pragma unroll
    for (int q = 0; q < C::kOwned; ++q) sq[q] = hist[q] = 0;

    const int32_t tile_s0 = p.s_begin + ts * p.tile_h;
    const int nbatch = p.tile_h / 4 + 4;
    // kept in a register: the asm compare takes its limit as a VGPR operand
    const uint32_t fifth_limit = __builtin_amdgcn_readfirstlane(kFifthLimit);

    for (int k = 0; k < nbatch; ++k) {
        const int row = 4 * k + wave;
        // ---- generation: this wave's slot row for all bands --------------
        {
            const uint32_t s_row = (uint32_t)tile_s0 - 8u + (uint32_t)row;
            uint32_t zhi[NB];
            uint32_t zlo = 0;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const uint32_t zc = (uint32_t)p.z0 + (uint32_t)b * (uint32_t)kPeriod + s_row;
                const uint64_t Z = z_term((int32_t)zc);
                zhi[b] = (uint32_t)(Z >> 16);
                if (b == 0) zlo = (uint32_t)Z << 16;
            }
            // The bit of chunk kChunksPerLane*l + j lands in bit j of each plane's
            // register, so the 64 lanes' bytes tile the 512-bit row.
            uint32_t acc[R::kPlanes];
#pragma unroll
            for (int q = 0; q < R::kPlanes; ++q) acc[q] = 0;
            static_for<0, kChunksPerLane>([&](auto jc) {
                constexpr int j = kChunksPerLane - 1 - decltype(jc)::value;  // high bits first
                const BandSlot sl =
                    band_slot(xlo[j], xh0[j], PH == 2 ? xh1[PH == 2 ? j : 0] : xh0[j], zlo, zhi[0]);
                acc[0] = acc[0] + acc[0] + (sl.e ? 1u : 0u);
                // all values first, so the multiplies of independent bands overlap
                uint32_t v[NB];
                static_for<0, NB>([&](auto bc) {
                    constexpr int b = decltype(bc)::value;
                    v[b] = band_value(sl, b, zhi[b]);
                });
                // shift in the complement of each hit: [a*K + bias > limit]
                static_for<0, NB>([&](auto bc) {
                    constexpr int b = decltype(bc)::value;
                    shift_in_above(acc[b + 1], v[b], fifth_limit);
                });
            });
            uint8_t* dst = reinterpret_cast<uint8_t*>(ring + (row % kRing) * R::kRowDw);
            dst[lane] = (uint8_t)acc[0];
#pragma unroll
            for (int q = 1; q < R::kPlanes; ++q) dst[q * 4 * kPlaneDw + lane] = (uint8_t)~acc[q];
        }
        __syncthreads();

        // ---- counting: rolling squares of the owned band-phases -----------
        if (counts) {
            uint32_t c4[C::kOwned];
#pragma unroll
            for (int q = 0; q < C::kOwned; ++q) c4[q] = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t* rowp = ring + ((4 * k + r) % kRing) * R::kRowDw;
                const uint32_t ew = window20(rowp, gbit);
                const uint32_t e0 = ~ew & 0xFFFFFu;  // slots with e == 0
#pragma unroll
                for (int hb = 0; hb < C::kBands; ++hb) {
                    const int b = b_lo + hb;
                    if (b >= NB) continue;  // odd NB: the second half is one band short
                    const uint32_t hw = window20(rowp + (b + 1) * kPlaneDw, gbit);
#pragma unroll
                    for (int pp = 0; pp < C::kPhases; ++pp) {
                        // phase ph of band b: where e ^ (b & 1) == ph
                        const bool use_e = ((b ^ (ph_lo + pp)) & 1) != 0;
                        c4[C::kPhases * hb + pp] += __popc(hw & (use_e ? ew : e0));
                    }
                }
            }
            uint32_t pass = 0;  // owned band-phases whose square reaches min_size
#pragma unroll
            for (int q = 0; q < C::kOwned; ++q) {
                const uint32_t square = (sq[q] & 0xFFFFu) + c4[q] - (sq[q] >> 16);
                sq[q] = square | ((hist[q] >> 24) << 16);
                hist[q] = (hist[q] << 8) | c4[q];
                pass |= (square >= (uint32_t)p.min_size ? 1u : 0u) << q;
            }
            const int m = k - 4;
            const int32_t s_g = tile_s0 + 4 * m;  // first slot row of the group
            if (m < 0 || col0 + gbit >= p.phase_cols || s_g >= p.s_end) pass = 0;
            while (pass) {
                const int q = __builtin_ctz(pass);
                pass &= pass - 1;
                const int b = b_lo + q / C::kPhases, ph = ph_lo + q % C::kPhases;
                if (b >= NB) continue;
                const int64_t ox0 = block_off + (int64_t)ph * kPeriod + col0 + gbit;
                if (ox0 >= p.w || ox0 + 3 < p.ox_min || (int64_t)b * kPeriod + s_g >= p.h_valid)
                    continue;
                refine_group<NB>(p, emit_min, ring, 4 * m, gbit, b, ph, ox0, col0 + gbit, s_g,
                                 lhist);
            }
        }
    }
}

// Adds the workgroup's LDS histogram to `out` and clears it.
__device__ __forceinline__ void flush_hist(uint32_t* lhist, unsigned long long* out) {
    __syncthreads();
    for (int i = threadIdx.x; i < kHistBins; i += kThreads) {
        const uint32_t v = lhist[i];
        if (v) {
            atomicAdd(&out[i], (unsigned long long)v);
            lhist[i] = 0;
        }
    }
}

}  // namespace band
}  // namespace slime
