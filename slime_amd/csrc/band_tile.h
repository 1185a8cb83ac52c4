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
// Counting: per band-phase a rolling 20x20 square of each 4x4 group. With the
// square bound and at least four band-phases, threads count plane words: the
// 20-row counts of 4-column blocks as bytes (nibble popcounts, 4-row strips,
// five strips kept), and the squares of the 8 groups whose windows start in
// the word as sums of 5 blocks in 16-bit lanes. Otherwise (few band-phases,
// or the union bound with its corner counts) a thread counts the 20-bit
// windows of its groups row by row. The groups that reach min_size are
// refined by the whole wave together (a bound
// on the union of the group's 16 disks, 20 lanes per group, then the exact
// 221-cell counts, 4 lanes per center), so a rare passing group costs the
// wave a few instructions instead of a long divergent run of one lane.
// Java's rejected first draws are not handled here; the engine recounts the
// few affected centers.
//
// With the union bound, two x-phases and 2 or 7-14 bands (the full world's
// launches and the commonest batched class), the 5-column window sums less
// the corner cells are int8 contractions on the matrix cores instead: the
// 20-row counts and corner counts of 16 band-phase rows x 32 nibble columns
// per wave are the A operand of v_mfma_i32_16x16x64_i8, a constant 0/+1/-1
// window matrix B, the threshold bias C; two instructions per wave and step
// give 16 rows x 32 groups, the second one with the counts of the next two
// plane words through LDS; above 8 bands a second 16-row operand set takes
// bands 8 and up. Build with SLIME_BAND_MFMA=0 to count those windows on the
// VALU.
#pragma once

#include <type_traits>

#include "scan.h"
#include "slime_math.h"

#ifndef SLIME_BAND_MFMA
#define SLIME_BAND_MFMA 1
#endif

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

// Per-nibble popcounts: nibble j of the result counts the set bits of nibble
// j of x.
__device__ __forceinline__ uint32_t nibble_counts(uint32_t x) {
    x = x - ((x >> 1) & 0x55555555u);
    return (x & 0x33333333u) + ((x >> 2) & 0x33333333u);
}

// v of the next lane in the lane's row of 16 (DPP row_shl:1); the last lane of
// a row reads 0.
__device__ __forceinline__ uint32_t from_next_lane16(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, true);
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

// Bitplanes of a ring row: E, one plane per band for x-phases 0 and 1, and
// for the phase without a partner (phase 0 with one phase, phase 2 with
// three) one plane per band pair (2k, 2k + 1), the last band alone when NB is
// odd.
template <int NB, int PH>
struct Ring {
    static constexpr int kPlanes = 1 + (PH >= 2 ? NB : 0) + (PH != 2 ? (NB + 1) / 2 : 0);
    static constexpr int kRowDw = kPlanes * kPlaneDw + 2;  // + pad for window reads
    static constexpr int kDwords = kRing * kRowDw;
};

// The plane of band b, phase ph: phases 0 and 1 share their band's plane (E
// tells which of them is the even one); the phase without a partner shares a
// plane per band pair (E tells which band). Either way band b, phase ph is
// the plane's bitmap under E where (b ^ ph) is odd and under ~E where it is
// even.
template <int NB, int PH>
__host__ __device__ constexpr int plane_of(int b, int ph) {
    return PH == 1 ? 1 + (b >> 1) : PH == 2 || ph < 2 ? 1 + b : 1 + NB + (b >> 1);
}

// LDS of one workgroup: the bitplane ring and the score histogram.
template <int NB, int PH>
struct TileLds {
    uint32_t ring[Ring<NB, PH>::kDwords];
    uint32_t lhist[kHistBins];
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

// Disk half-widths 8,8,8,7,7,6,5,4,2 for |dz| = 0..8 as nibbles.
constexpr uint64_t kHalfWidthNibbles = 0x245677888ull;
constexpr int half_width_of(int adz) { return (int)((kHalfWidthNibbles >> (4 * adz)) & 15u); }

// Half-width of the union of the 16 disks of a 4x4 group in group row r
// (0..19): the widest disk row any of the group's 4 center rows reaches.
constexpr int union_half_width(int r) {
    const int d = r < 19 - r ? r : 19 - r;
    return d >= 6 ? 8 : d == 0 ? 2 : d >= 4 ? 7 : d + 3;
}

// Both agree with the disk definition (dx^2 + dz^2 <= 68).
constexpr int disk_half_width_ref(int adz) {
    int h = 0;
    while ((h + 1) * (h + 1) + adz * adz <= 68) ++h;
    return h;
}
constexpr bool check_tables() {
    for (int adz = 0; adz <= 8; ++adz)
        if (half_width_of(adz) != disk_half_width_ref(adz)) return false;
    for (int r = 0; r < 20; ++r) {
        int h = -1;
        for (int j = 0; j < 4; ++j) {
            const int dz = r - 8 - j;
            if (dz >= -8 && dz <= 8) {
                const int w = disk_half_width_ref(dz < 0 ? -dz : dz);
                h = w > h ? w : h;
            }
        }
        if (h != union_half_width(r)) return false;
    }
    return true;
}
static_assert(check_tables(), "disk and union row tables");

// Row r of a group's union of disks: bits [8 - h, 11 + h] of the window.
__host__ __device__ constexpr uint32_t union_row_mask(int r) {
    return ((1u << (4 + 2 * union_half_width(r))) - 1u) << (8 - union_half_width(r));
}

// The window bits of row r outside the union (non-zero for r <= 5, r >= 14;
// the rolling bound trims rows 0..3 and 16..19, 60 of the 68 cells).
__host__ __device__ constexpr uint32_t corner_mask(int r) { return ~union_row_mask(r) & 0xFFFFFu; }
static_assert(corner_mask(6) == 0 && corner_mask(13) == 0 && corner_mask(5) != 0 &&
                  corner_mask(14) != 0,
              "corner rows");

// A group of one band-phase, as a lane hands it to its wave:
// (gbit << 8) | (b << 2) | ph, gbit = 4 * group.
__device__ __forceinline__ uint32_t group_code(int gbit, int b, int ph) {
    return ((uint32_t)gbit << 8) | ((uint32_t)b << 2) | (uint32_t)ph;
}

typedef int mfma_i8x16 __attribute__((ext_vector_type(4)));  // 16 int8 / 4 int32

// Matrix-core window sums with two x-phases and NB (2..14) bands: QN = 2 NB
// band-phases. Up to 8 bands, QN band-phases and H column ranges make the 16
// rows of A (H QN <= 16 rows used); lane l of counting wave w holds row
// m = l % 16 (band-phase m % QN, range m / QN) over plane word
// (m / QN) * kRangeWords + 4 w + l / 16, i.e. a wave covers 32 nibble columns
// of each range; the waves past kWaves do not count. Above 8 bands there are
// two 16-row sets (kSets): lane l holds row m of set s, band-phase 16 s + m
// (set 1 has QN - 16 rows), both over plane word 4 w + l / 16. D (lane l,
// register i): row 4 (l / 16) + i, group l % 16.
template <int NB>
struct MfmaRoles {
    static constexpr int QN = 2 * NB;
    static constexpr int kSets = QN > 16 ? 2 : 1;
    static constexpr int H = kSets == 2 || QN > 8 ? 1 : QN > 4 ? 2 : 4;
    static constexpr int kRows = kSets == 2 ? 16 : QN * H;  // rows of set 0
    static constexpr int kRows1 = kSets == 2 ? QN - 16 : 0;  // rows of set 1
    static constexpr int kRangeWords = 16 / H;
    static constexpr int kWaves = 4 / H;
    static_assert(NB >= 2 && NB <= 16 && kRows <= 16 && kRows1 <= 16, "2..16 bands");
};

// B of the window contraction for lane l: column n = l % 16 (the n-th group
// of the operand's first plane word), k-block l / 16 (its (l / 16)-th plane
// word): bytes 0-3 the word's even nibble columns' counts, 4-7 the odd ones,
// 8-11 / 12-15 their corner counts. Group n's window: nibble columns n..n+4,
// less the corner counts of columns n and n+4.
__device__ __forceinline__ mfma_i8x16 window_matrix(int lane) {
    const int n = lane & 15, j = lane >> 4;
    int dw[4];
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
        dw[d] = (int)word;
    }
    return mfma_i8x16{dw[0], dw[1], dw[2], dw[3]};
}

// Row r (0..19) of the 20-bit window of a coded group: slime chunks of its
// band-phase starting at group row `first` of the ring.
template <int NB, int PH>
__device__ __forceinline__ uint32_t group_row(const uint32_t* ring, int first, uint32_t code,
                                              int r) {
    const int gbit = (int)(code >> 8), b = (int)((code >> 2) & 15u);
    const int ph = (int)(code & (PH == 3 ? 3u : 1u));
    const uint32_t flip = ((b ^ ph) & 1) ? 0u : 0xFFFFFFFFu;  // phase mask is E or ~E
    // first >= 0 wherever a group is refined: an unsigned remainder is cheaper
    const uint32_t* row =
        ring + ((uint32_t)(first + r) % (uint32_t)kRing) * Ring<NB, PH>::kRowDw;
    return window20(row + plane_of<NB, PH>(b, ph) * kPlaneDw, gbit) &
           (window20(row, gbit) ^ flip);
}

// The wave refines the groups its lanes found (todo: a bit per owned
// band-phase q, coded by code_of(q)). All 64 lanes must be active.
//  1. (square bound, UB false) Up to three groups at a time: lanes
//     20s..20s+19 count row r of group s under the union-of-disks mask; the
//     five bits of each count are summed per group with ballots (a scalar
//     popcount each).
//  2. A group whose union count reaches min_size (with UB, every group the
//     lanes found): exact counts (exact_group).
// Exact 221-cell counts of the 16 centers of one coded group by the whole
// wave: lane r < 20 reads window row r once; lane 4c + part takes the disk
// rows dz = -8 + part + 4t of center c from those lanes (ds_bpermute), a quad
// reduction gives the count, and lane 4c emits it.
template <int NB, int PH, typename EmitCenter>
__device__ __forceinline__ void exact_group(uint32_t code, const uint32_t* ring, int first,
                                            const EmitCenter& emit) {
    const int lane = threadIdx.x & 63;
    const uint32_t row = group_row<NB, PH>(ring, first, code, lane < 20 ? lane : 19);
    const int center = lane >> 2, part = lane & 3;
    const int i = center & 3, jr = center >> 2;
    uint32_t cnt = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        // dz <= 8 except for t = 4, part > 0; every lane takes part in the
        // shuffle (a lane outside it would read as zero)
        const int dz = -8 + part + 4 * t;
        const uint32_t w = (uint32_t)__shfl((int)row, 8 + jr + (dz <= 8 ? dz : 8));
        if (dz <= 8) {
            const int hw = half_width_of(dz < 0 ? -dz : dz);
            cnt += __popc(w & (((1u << (2 * hw + 1)) - 1u) << (8 + i - hw)));
        }
    }
    // quad sums by DPP quad permutes: lanes 1,0,3,2 then 2,3,0,1
    cnt += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)cnt, 0xB1, 0xF, 0xF, false);
    cnt += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)cnt, 0x4E, 0xF, 0xF, false);
    if (part == 0) emit(code, i, jr, cnt);
}

template <int NB, int PH, bool UB, typename CodeOf, typename EmitCenter>
__device__ __forceinline__ void refine_wave(uint32_t todo, const CodeOf& code_of,
                                            const uint32_t* ring, int first, int32_t min_size,
                                            const EmitCenter& emit) {
    if constexpr (UB) {
        // the lanes' union bounds already passed: exact counts, a group at a time
        for (;;) {
            const uint64_t act = __ballot(todo != 0);
            if (act == 0) break;
            const int l = __builtin_ctzll(act);
            const uint32_t mine = todo ? code_of(__builtin_ctz(todo)) : 0u;
            const uint32_t code = __builtin_amdgcn_readlane(mine, l);
            if ((threadIdx.x & 63) == l) todo &= todo - 1;
            exact_group<NB, PH>(code, ring, first, emit);
        }
        return;
    }
    const int lane = threadIdx.x & 63;
    const int slot = lane < 60 ? lane / 20 : 3;
    const int r = lane - 20 * slot;
    for (;;) {
        uint64_t act = __ballot(todo != 0);
        if (act == 0) break;
        const uint32_t mine = todo ? code_of(__builtin_ctz(todo)) : 0u;
        uint32_t code[3] = {0u, 0u, 0u};
        int nsel = 0;
        uint64_t taken = 0;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            if (act != 0) {
                const int l = __builtin_ctzll(act);
                act &= act - 1;
                taken |= 1ull << l;
                code[s] = __builtin_amdgcn_readlane(mine, l);
                nsel = s + 1;
            }
        }
        if ((taken >> lane) & 1ull) todo &= todo - 1;
        const uint32_t my_code = slot == 0 ? code[0] : slot == 1 ? code[1] : code[2];
        uint32_t c = 0;
        if (slot < nsel)
            c = __popc(group_row<NB, PH>(ring, first, my_code, r) & union_row_mask(r));
        uint64_t bits[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) bits[j] = __ballot((c >> j) & 1u);
        for (int s = 0; s < nsel; ++s) {
            const uint64_t m = 0xFFFFFull << (20 * s);
            uint32_t sum = 0;
#pragma unroll
            for (int j = 0; j < 5; ++j) sum += (uint32_t)__popcll(bits[j] & m) << j;
            if (sum < (uint32_t)min_size) continue;
            exact_group<NB, PH>(code[s], ring, first, emit);
        }
    }
}

// Scans tile `tile` of the job with the workgroup's LDS; scores accumulate in
// lds.lhist. The caller synchronises the workgroup before the call when the
// LDS was in use.
// UB selects the bound that sends a group to refinement: false, the rolling
// 20x20 square (then the union of disks is counted in refine_wave); true,
// the square minus the corner cells (corner_mask) of the window's first and
// last four rows, 340 cells that hold the union of disks. A new row's corner
// count at offset 16 + r is taken now and at offset r (four steps later) is
// queued a byte per step. Two masked popcounts per row and band-phase, and no
// refinement of the many groups whose square passes at thresholds near 50.
// tid is the thread's index in the workgroup (threadIdx.x; a loop over tiles
// passes a copy the compiler cannot see through, so that the tile's per-lane
// constants are computed per tile rather than kept live across the loop).
template <int NB, int PH, bool UB>
__device__ __forceinline__ void band_tile(const BandParams& p, int32_t emit_min, int tile,
                                          TileLds<NB, PH>& lds, int tid) {
    static_assert(PH >= 1 && PH <= 3, "one to three x-phases");
    // three phases: 2..5 bands, the union bound, counting by plane words
    // with a thread group per band-phase
    static_assert(PH < 3 || (NB >= 2 && NB <= 5 && UB), "three-phase classes");
    using R = Ring<NB, PH>;
    using C = Roles<NB, PH>;
    uint32_t* ring = lds.ring;
    uint32_t* lhist = lds.lhist;
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
    // center offset of phase 0 (x-blocks of PH P columns; one block with PH = 1)
    const int64_t block_off = (int64_t)xblock * (PH == 1 ? 2 : PH) * kPeriod;
    const int32_t col0 = p.slot_lo + xtile * kBandTileOut;      // slot column offset in phase
    const uint32_t chunk_x0 = (uint32_t)p.x0 + (uint32_t)(block_off + col0) - 8u;

    // Lane l evaluates the chunks kChunksPerLane*l + j of every row.
    uint32_t xlo[kChunksPerLane], xh0[kChunksPerLane], xh1[PH >= 2 ? kChunksPerLane : 1];
    uint32_t xh2[PH == 3 ? kChunksPerLane : 1];
#pragma unroll
    for (int j = 0; j < kChunksPerLane; ++j) {
        const uint32_t xc = chunk_x0 + (uint32_t)(kChunksPerLane * lane + j);
        const uint64_t X0 = (uint64_t)p.seed + x_term((int32_t)xc);
        xlo[j] = (uint32_t)X0 << 16;
        xh0[j] = (uint32_t)(X0 >> 16);
        if constexpr (PH >= 2) {
            const uint64_t X1 = (uint64_t)p.seed + x_term((int32_t)(xc + (uint32_t)kPeriod));
            xh1[j] = (uint32_t)(X1 >> 16);
        }
        if constexpr (PH == 3) {
            const uint64_t X2 = (uint64_t)p.seed + x_term((int32_t)(xc + 2u * (uint32_t)kPeriod));
            xh2[j] = (uint32_t)(X2 >> 16);
        }
    }

    // Counting role: group g, bands [b_lo, b_lo + kBands), phases [ph_lo, +kPhases).
    const int g = C::kPerGroup == 2 ? tid >> 1 : tid;
    const int half = C::kPerGroup == 2 ? (tid & 1) : 0;
    const int b_lo = C::kSplitPhase ? 0 : half * C::kBands;
    const int ph_lo = C::kSplitPhase ? half : 0;
    const int gbit = 4 * g;
    const bool counts = g < kGroupCols;
    // E (not ~E) selects phase ph_lo of band b_lo when b_lo ^ ph_lo is odd
    const uint32_t flip0 = ((b_lo ^ ph_lo) & 1) ? 0u : 0xFFFFFu;
    // Rolling state per owned band-phase q = kPhases*(b - b_lo) + (ph - ph_lo):
    // sq = square | (strip leaving next) << 16; hist = last four strips, a byte each.
    uint32_t sq[C::kOwned], hist[C::kOwned];
    // UB: corner counts due 1..4 steps ahead, a byte each (byte 0 next step)
    uint32_t pend[UB ? C::kOwned : 1];
#pragma unroll
    for (int q = 0; q < C::kOwned; ++q) sq[q] = hist[q] = 0;
#pragma unroll
    for (int q = 0; q < (UB ? C::kOwned : 1); ++q) pend[q] = 0;

    // Square bound with at least four band-phases: counting by plane words
    // (fewer instructions per band than by groups; with fewer band-phases the
    // work would sit on too few threads). Thread t takes band
    // wb = t / 16 over plane word wd = t % 16, i.e. the 8 nibble columns
    // 8 wd .. 8 wd + 7 of the row and the windows (5 nibble columns each) of
    // the groups 8 wd + i, for all its phases. Per phase it keeps the 20-row
    // counts of its nibble columns as bytes (even columns in `we`, odd in
    // `wo`, at most 80 each) and its last five 4-row strips in the same form;
    // the next word's counts come from the next lane of the row of 16.
    // With two phases and at most 8 bands a thread takes one phase (wph) of
    // its band, so that 16 x NB x 2 threads share the work.
    constexpr bool kWord = NB * PH >= 4;
    // With the union bound, word counting takes a bound between the square and
    // the union of disks: the square minus its first and last 4-column block
    // in window rows 0, 1, 18 and 19 (32 cells, all outside the union; 368
    // remain). Groups that pass it are refined as with the square bound.
    constexpr bool kCorner = UB && kWord;
    // a thread group of 16 per band-phase: two phases and up to 8 bands, or
    // three phases
    constexpr bool kWSplit = (PH == 2 && NB <= 8) || PH == 3;
    constexpr int kWThreads = 16 * NB * (kWSplit ? PH : 1);  // threads that hold a band
    // Matrix-core window sums (MfmaRoles): the lane's plane word, band and
    // phase follow its row of A instead.
    // (measured per class: faster with 2 and 7-8 bands, and the only
    // union-bound form above 8; with 3-6 bands the exchange's workgroup
    // barrier, or a fourth counting wave where the VALU needs three, costs
    // more than the window sums save)
    constexpr bool kMfma = SLIME_BAND_MFMA && kCorner && PH == 2 && (NB == 2 || NB >= 7);
    using MR = MfmaRoles<kMfma ? NB : 2>;
    const int mrow = lane & 15;
    const bool m_ok = kMfma && wave < MR::kWaves && mrow < MR::kRows;
    const int wd = !kMfma ? (tid & 15)
                 : m_ok   ? (mrow / MR::QN) * MR::kRangeWords + 4 * wave + (lane >> 4)
                          : 0;
    // three phases: band-phase t / 16 is (band, phase 0 or 1) below 2 NB,
    // then (band, phase 2)
    const int wbp = tid >> 4;
    const int wb = PH == 3 ? (wbp < 2 * NB ? wbp >> 1 : wbp - 2 * NB)
                 : !kMfma  ? (kWSplit ? tid >> 5 : tid >> 4)
                 : m_ok    ? (mrow % MR::QN) >> 1
                           : NB;
    const int wph = PH == 3 ? (wbp < 2 * NB ? wbp & 1 : 2)
                  : !kMfma  ? (kWSplit ? (tid >> 4) & 1 : 0)
                            : (mrow % MR::QN) & 1;
    const bool w_active = wb < NB;
    const int wbc = w_active ? wb : 0;  // idle threads read a valid plane
    // phase 0 of band b lies where e == (b & 1): E for odd b, ~E for even b;
    // a thread of phase 1 counts the rest
    const uint32_t wflip = ((wbc ^ wph) & 1) ? 0u : 0xFFFFFFFFu;
    constexpr int kWPh = !kWord || kWSplit ? 1 : PH;
    uint32_t we[kWPh], wo[kWPh], he[kWPh][5], ho[kWPh][5];
    // kCorner: per-nibble counts of rows 0 and 1 of the last four strips
    // (at most 8 a nibble), [0] oldest
    uint32_t hq[kCorner ? kWPh : 1][4];
#pragma unroll
    for (int ph = 0; ph < (kCorner ? kWPh : 1); ++ph)
#pragma unroll
        for (int j = 0; j < 4; ++j) hq[ph][j] = 0;
    uint32_t wvalid[kWPh];  // bit i: group 8 wd + i has a center inside the job
#pragma unroll
    for (int ph = 0; ph < kWPh; ++ph) {
        we[ph] = wo[ph] = 0;
#pragma unroll
        for (int j = 0; j < 5; ++j) he[ph][j] = ho[ph][j] = 0;
        wvalid[ph] = 0;
        if constexpr (kWord && !kMfma) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int gi = 8 * wd + i;
                const int64_t ox0 = block_off + (int64_t)(ph + wph) * kPeriod + col0 + 4 * gi;
                if (gi < kGroupCols && col0 + 4 * gi < p.phase_cols && ox0 < p.w &&
                    ox0 + 3 >= p.ox_min)
                    wvalid[ph] |= 1u << i;
            }
        }
    }
    // 0x8000 - min_size in both 16-bit lanes: bit 15 / 31 of square + this is
    // set when the lane's square reaches min_size
    const uint32_t wbias = (uint32_t)(0x8000 - p.min_size) * 0x10001u;
    // kMfma: bit 8 s + 4 f + i of `mvalid` says that D register i of the f-th
    // contraction of set s (row 4 (lane / 16) + i, group 16 f + lane % 16 of
    // the row's first plane word) has a center inside the job; rows i = 0, 1
    // are band mband[s][0], rows 2, 3 band mband[s][1]. C is 0x8000 -
    // min_size, so that bit 15 of a sum says it reaches min_size. Each set has
    // its exchange (18 words of 16 band-phases); words 16 and 17 (past the
    // row) stay zero. With two sets, set 1's band is 8 more than set 0's (the
    // same phase and E / ~E selection); the lanes past its rows read set 0's
    // plane and give zero rows.
    constexpr int kXchSet = (kPlaneDw + 2) * 16;
    __shared__ mfma_i8x16 mxch[kMfma ? kXchSet * MR::kSets : 1];
    const bool m_ok1 = kMfma && MR::kSets == 2 && wave < MR::kWaves && mrow < MR::kRows1;
    const int mplane1 = m_ok1 ? wbc + 9 : wbc + 1;
    const int mq = kMfma && MR::kSets == 1 ? mrow % MR::QN : mrow;  // row within its set
    uint32_t mvalid = 0;
    int mband[MR::kSets][2];
#pragma unroll
    for (int s = 0; s < MR::kSets; ++s) mband[s][0] = mband[s][1] = NB;
    mfma_i8x16 mB = {0, 0, 0, 0}, mC = {0, 0, 0, 0};
    if constexpr (kMfma) {
        static_assert(MR::kSets == kWPh, "one rolling state per operand set");
        if (tid < 32 * MR::kSets)
            mxch[(tid >> 5) * kXchSet + kPlaneDw * 16 + (tid & 31)] = mfma_i8x16{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < MR::kSets; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int md = 4 * (lane >> 4) + i;
                if (wave >= MR::kWaves) continue;
                int qd, hd;
                if constexpr (MR::kSets == 2) {
                    qd = 16 * s + md;
                    hd = 0;
                    if (qd >= MR::QN) continue;
                } else {
                    if (md >= MR::kRows) continue;
                    qd = md % MR::QN;
                    hd = md / MR::QN;
                }
                if ((i & 1) == 0) mband[s][i >> 1] = qd >> 1;
#pragma unroll
                for (int f = 0; f < 2; ++f) {
                    const int g = 8 * (hd * MR::kRangeWords + 4 * wave) + 16 * f + (lane & 15);
                    const int64_t ox0 = block_off + (int64_t)(qd & 1) * kPeriod + col0 + 4 * g;
                    if (g < kGroupCols && col0 + 4 * g < p.phase_cols && ox0 < p.w &&
                        ox0 + 3 >= p.ox_min)
                        mvalid |= 1u << (8 * s + 4 * f + i);
                }
            }
        mB = window_matrix(lane);
        const int bias = 0x8000 - p.min_size;
        mC = mfma_i8x16{bias, bias, bias, bias};
    }

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
            // With two phases and 12 or more bands the carry out of the slot's
            // low half is added once per slot to each of the two selected phase
            // words, and each band's sum is then a two-operand add of VGPRs (the
            // band's z word copied to a VGPR once per row): a single-rate add
            // instead of an add with carry-in (half rate). Measured per class
            // (kernel T centers/s): 12-14 bands 0.5-0.9% faster, 4-8 bands
            // 1-3% slower (the two adds per slot and their registers), so only
            // the wide classes take it.
            constexpr bool kPreCarry = PH == 2 && NB >= 12;
            uint32_t zv[kPreCarry ? NB : 1];
            if constexpr (kPreCarry) {
#pragma unroll
                for (int b = 0; b < NB; ++b) asm("v_mov_b32 %0, %1" : "=v"(zv[b]) : "s"(zhi[b]));
            }
            // The z words of the band pairs tested as one (the phase without a
            // partner) copied to VGPRs once per row, so that each slot selects
            // its word with one v_cndmask: selected from the scalar words, the
            // choice of two neighbouring array elements became an indexed LDS
            // read per slot.
            constexpr int kZPairs = PH != 2 ? NB / 2 : 0;
            uint32_t zpe[kZPairs > 0 ? kZPairs : 1], zpo[kZPairs > 0 ? kZPairs : 1];
#pragma unroll
            for (int k = 0; k < kZPairs; ++k) {
                asm("v_mov_b32 %0, %1" : "=v"(zpe[k]) : "s"(zhi[2 * k]));
                asm("v_mov_b32 %0, %1" : "=v"(zpo[k]) : "s"(zhi[2 * k + 1]));
            }
            // The bit of chunk kChunksPerLane*l + j lands in bit j of each plane's
            // register, so the 64 lanes' bytes tile the 512-bit row.
            uint32_t acc[R::kPlanes];
#pragma unroll
            for (int q = 0; q < R::kPlanes; ++q) acc[q] = 0;
            static_for<0, kChunksPerLane>([&](auto jc) {
                constexpr int j = kChunksPerLane - 1 - decltype(jc)::value;  // high bits first
                const BandSlot sl =
                    band_slot(xlo[j], xh0[j], PH >= 2 ? xh1[PH >= 2 ? j : 0] : xh0[j], zlo, zhi[0]);
                acc[0] = acc[0] + acc[0] + (sl.e ? 1u : 0u);
                // all values first, so the multiplies of independent bands overlap
                constexpr int kXPairs = PH >= 2 ? NB : 0;  // tests of x-phase pairs
                uint32_t v[PH >= 2 ? NB : 1];
                if constexpr (PH == 1) {
                } else if constexpr (kPreCarry) {
                    uint32_t xa = sl.xa + sl.carry, xb = sl.xb + sl.carry;
                    asm("" : "+v"(xa), "+v"(xb));  // the bands add to these sums
                    static_for<0, NB>([&](auto bc) {
                        constexpr int b = decltype(bc)::value;
                        v[b] = ((((b & 1) ? xb : xa) + zv[b]) ^ kHighScramble) * kMulInv5 + sl.bias;
                    });
                } else {
                    // with one band, the bias as a value of its own (as for
                    // one phase below: one multiply-add per test)
                    BandSlot sb = sl;
                    if constexpr (NB == 1) asm("" : "+v"(sb.bias));
                    static_for<0, NB>([&](auto bc) {
                        constexpr int b = decltype(bc)::value;
                        v[b] = band_value(sb, b, zhi[b]);
                    });
                }
                // shift in the complement of each hit: [a*K + bias > limit]
                static_for<0, kXPairs>([&](auto bc) {
                    constexpr int b = decltype(bc)::value;
                    shift_in_above(acc[b + 1], v[b], fifth_limit);
                });
                if constexpr (PH != 2) {
                    // the phase without a partner (phase 0 with one phase;
                    // phase 2 = x + 2P, which has phase 0's parity, with
                    // three): of the band pair (2k, 2k + 1) the even chunk is
                    // band 2k's where e == 0 and band 2k + 1's where e == 1 (an
                    // odd NB's last band is tested alone; the plane's mask
                    // keeps it where it is even)
                    constexpr int kPairs = (NB + 1) / 2;
                    const uint32_t xs = PH == 1 ? xh0[j] : xh2[PH == 3 ? j : 0];
                    // the slot's bias as a value of its own: otherwise the
                    // compiler folds its product into each test's sum (two
                    // multiplies and a three-operand add where a multiply-add
                    // does)
                    uint32_t bias = sl.bias;
                    if constexpr (PH == 1) asm("" : "+v"(bias));
                    uint32_t v2[kPairs];
                    static_for<0, kPairs>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        uint32_t zs = zhi[2 * k];
                        if constexpr (2 * k + 1 < NB) zs = sl.e ? zpo[k] : zpe[k];
                        v2[k] = ((xs + zs + sl.carry) ^ kHighScramble) * kMulInv5 + bias;
                    });
                    static_for<0, kPairs>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        shift_in_above(acc[1 + kXPairs + k], v2[k], fifth_limit);
                    });
                }
            });
            uint8_t* dst = reinterpret_cast<uint8_t*>(ring + (row % kRing) * R::kRowDw);
            dst[lane] = (uint8_t)acc[0];
#pragma unroll
            for (int q = 1; q < R::kPlanes; ++q) dst[q * 4 * kPlaneDw + lane] = (uint8_t)~acc[q];
        }
        __syncthreads();

        // ---- counting: rolling squares of the owned band-phases -----------
        const int m = k - 4;
        const int32_t s_g = tile_s0 + 4 * m;  // first slot row of the groups
        uint32_t pass = 0;  // owned band-phases whose square reaches min_size
        if constexpr (kMfma) {
            // the vertical part as below (one phase a thread, per operand
            // set), then the lane's counts and corner counts as its 16 bytes
            // of A
            mfma_i8x16 a0[MR::kSets];
#pragma unroll
            for (int s = 0; s < MR::kSets; ++s) a0[s] = mfma_i8x16{0, 0, 0, 0};
            if (wave < MR::kWaves) {
                uint32_t nhead[MR::kSets], ntail[MR::kSets];
#pragma unroll
                for (int s = 0; s < MR::kSets; ++s) nhead[s] = ntail[s] = 0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t* rowp = ring + ((4 * k + r) % kRing) * R::kRowDw;
                    const uint32_t sel = rowp[wd] ^ wflip;
#pragma unroll
                    for (int s = 0; s < MR::kSets; ++s) {
                        const int plane = s == 0 ? wbc + 1 : mplane1;
                        const uint32_t c = nibble_counts(rowp[plane * kPlaneDw + wd] & sel);
                        if (r < 2) nhead[s] += c;
                        else ntail[s] += c;
                    }
                }
#pragma unroll
                for (int s = 0; s < MR::kSets; ++s) {
                    const uint32_t se = (nhead[s] & 0x0F0F0F0Fu) + (ntail[s] & 0x0F0F0F0Fu);
                    const uint32_t so =
                        ((nhead[s] >> 4) & 0x0F0F0F0Fu) + ((ntail[s] >> 4) & 0x0F0F0F0Fu);
                    we[s] += se - he[s][0];
                    wo[s] += so - ho[s][0];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        he[s][j] = he[s][j + 1];
                        ho[s][j] = ho[s][j + 1];
                    }
                    he[s][4] = se;
                    ho[s][4] = so;
                    const uint32_t old = hq[s][0];
#pragma unroll
                    for (int j = 0; j < 3; ++j) hq[s][j] = hq[s][j + 1];
                    hq[s][3] = nhead[s];
                    const uint32_t pe = (ntail[s] & 0x0F0F0F0Fu) + (old & 0x0F0F0F0Fu);
                    const uint32_t po =
                        ((ntail[s] >> 4) & 0x0F0F0F0Fu) + ((old >> 4) & 0x0F0F0F0Fu);
                    if (s == 0 ? m_ok : m_ok1) {
                        a0[s] = mfma_i8x16{(int)we[s], (int)wo[s], (int)pe, (int)po};
                        mxch[s * kXchSet + wd * 16 + mq] = a0[s];
                    }
                }
            }
            if constexpr (MR::kWaves == 1) {
                // one counting wave: its own lanes wrote the words it reads
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else {
                __syncthreads();
            }
            if (wave < MR::kWaves) {
                // the second contraction starts two plane words on; the
                // exchange is indexed by band-phase, since with several column
                // ranges the rows of another range hold the word
#pragma unroll
                for (int s = 0; s < MR::kSets; ++s) {
                    mfma_i8x16 a1 = {0, 0, 0, 0};
                    if (s == 0 ? m_ok : m_ok1) a1 = mxch[s * kXchSet + (wd + 2) * 16 + mq];
                    const mfma_i8x16 d0 =
                        __builtin_amdgcn_mfma_i32_16x16x64_i8(a0[s], mB, mC, 0, 0, 0);
                    const mfma_i8x16 d1 =
                        __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, mB, mC, 0, 0, 0);
                    const uint32_t lo = (((uint32_t)d0[0] & 0x8000u) |
                                         (((uint32_t)d0[1] & 0x8000u) << 1) |
                                         (((uint32_t)d0[2] & 0x8000u) << 2) |
                                         (((uint32_t)d0[3] & 0x8000u) << 3)) >> 15;
                    const uint32_t hi = (((uint32_t)d1[0] & 0x8000u) |
                                         (((uint32_t)d1[1] & 0x8000u) << 1) |
                                         (((uint32_t)d1[2] & 0x8000u) << 2) |
                                         (((uint32_t)d1[3] & 0x8000u) << 3)) >> 11;
                    pass |= (lo | hi) << (8 * s);
                }
                pass &= mvalid;
                if (m < 0 || s_g >= p.s_end) pass = 0;
                if (pass) {
#pragma unroll
                    for (int s = 0; s < MR::kSets; ++s) {
                        if ((int64_t)mband[s][0] * kPeriod + s_g >= p.h_valid)
                            pass &= ~(0x33u << (8 * s));
                        if ((int64_t)mband[s][1] * kPeriod + s_g >= p.h_valid)
                            pass &= ~(0xCCu << (8 * s));
                    }
                }
            }
        } else if constexpr (kWord) {
            // every thread of a wave that holds a band takes part (the next-lane
            // reads need whole rows); the waves past the last band skip it
            if (wave * 64 < kWThreads) {
                uint32_t nhead[kWPh], ntail[kWPh];  // rows 0-1 and 2-3 of the strip
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t* rowp = ring + ((4 * k + r) % kRing) * R::kRowDw;
                    const uint32_t e = rowp[wd];
                    const uint32_t h = rowp[plane_of<NB, PH>(wbc, wph) * kPlaneDw + wd];
                    uint32_t c[kWPh];
                    c[0] = nibble_counts(h & (e ^ wflip));
                    if constexpr (kWPh == 2) c[1] = nibble_counts(h) - c[0];  // the rest
#pragma unroll
                    for (int ph = 0; ph < kWPh; ++ph) {  // two rows: at most 8 a nibble
                        if (r == 0) nhead[ph] = c[ph];
                        else if (r == 1) nhead[ph] += c[ph];
                        else if (r == 2) ntail[ph] = c[ph];
                        else ntail[ph] += c[ph];
                    }
                }
#pragma unroll
                for (int ph = 0; ph < kWPh; ++ph) {
                    // the strip's counts as bytes: even nibble columns, odd ones
                    const uint32_t se = (nhead[ph] & 0x0F0F0F0Fu) + (ntail[ph] & 0x0F0F0F0Fu);
                    const uint32_t so =
                        ((nhead[ph] >> 4) & 0x0F0F0F0Fu) + ((ntail[ph] >> 4) & 0x0F0F0F0Fu);
                    we[ph] += se - he[ph][0];
                    wo[ph] += so - ho[ph][0];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        he[ph][j] = he[ph][j + 1];
                        ho[ph][j] = ho[ph][j + 1];
                    }
                    he[ph][4] = se;
                    ho[ph][4] = so;
                    // nibble columns n = 8 wd + j as 16-bit lanes: W_s holds
                    // (n_s, n_{s+4}), N_s the next word's (n_{s+8}, n_{s+12})
                    const uint32_t ne = from_next_lane16(we[ph]), no = from_next_lane16(wo[ph]);
                    const uint32_t w0 = we[ph] & 0x00FF00FFu, w2 = (we[ph] >> 8) & 0x00FF00FFu;
                    const uint32_t w1 = wo[ph] & 0x00FF00FFu, w3 = (wo[ph] >> 8) & 0x00FF00FFu;
                    const uint32_t n0 = ne & 0x00FF00FFu, n2 = (ne >> 8) & 0x00FF00FFu;
                    const uint32_t n1 = no & 0x00FF00FFu, n3 = (no >> 8) & 0x00FF00FFu;
                    // square of group s (lane 0) and s + 4 (lane 1): 5 nibble columns
                    uint32_t q0 = w0 + w1 + w2 + w3 + __builtin_amdgcn_alignbit(n0, w0, 16);
                    uint32_t q1 = q0 - w0 + __builtin_amdgcn_alignbit(n1, w1, 16);
                    uint32_t q2 = q1 - w1 + __builtin_amdgcn_alignbit(n2, w2, 16);
                    uint32_t q3 = q2 - w2 + __builtin_amdgcn_alignbit(n3, w3, 16);
                    if constexpr (kCorner) {
                        // window rows 18, 19 (this strip's rows 2, 3) and 0, 1 (rows
                        // 0, 1 of the strip four steps back), per nibble column as
                        // bytes; a group loses its first and last column's counts
                        const uint32_t old = hq[ph][0];
#pragma unroll
                        for (int j = 0; j < 3; ++j) hq[ph][j] = hq[ph][j + 1];
                        hq[ph][3] = nhead[ph];
                        const uint32_t pe = (ntail[ph] & 0x0F0F0F0Fu) + (old & 0x0F0F0F0Fu);
                        const uint32_t po =
                            ((ntail[ph] >> 4) & 0x0F0F0F0Fu) + ((old >> 4) & 0x0F0F0F0Fu);
                        const uint32_t npe = from_next_lane16(pe), npo = from_next_lane16(po);
                        const uint32_t p0 = pe & 0x00FF00FFu, p2 = (pe >> 8) & 0x00FF00FFu;
                        const uint32_t p1 = po & 0x00FF00FFu, p3 = (po >> 8) & 0x00FF00FFu;
                        q0 -= p0 + __builtin_amdgcn_alignbit(npe & 0x00FF00FFu, p0, 16);
                        q1 -= p1 + __builtin_amdgcn_alignbit(npo & 0x00FF00FFu, p1, 16);
                        q2 -= p2 + __builtin_amdgcn_alignbit((npe >> 8) & 0x00FF00FFu, p2, 16);
                        q3 -= p3 + __builtin_amdgcn_alignbit((npo >> 8) & 0x00FF00FFu, p3, 16);
                    }
                    const uint32_t c0 = q0 + wbias, c1 = q1 + wbias, c2 = q2 + wbias,
                                   c3 = q3 + wbias;
                    const uint32_t pm =
                        ((c0 >> 15) & 1u) | ((c1 >> 14) & 2u) | ((c2 >> 13) & 4u) |
                        ((c3 >> 12) & 8u) | ((c0 >> 27) & 16u) | ((c1 >> 26) & 32u) |
                        ((c2 >> 25) & 64u) | ((c3 >> 24) & 128u);
                    pass |= (pm & wvalid[ph]) << (8 * ph);
                }
                if (!w_active || m < 0 || s_g >= p.s_end ||
                    (int64_t)wb * kPeriod + s_g >= p.h_valid)
                    pass = 0;
            }
        } else if (counts) {
            uint32_t c4[C::kOwned];
            // UB: this strip's corner cells at offsets 16+r (now) and r (in
            // four steps)
            uint32_t k0[UB ? C::kOwned : 1], k4[UB ? C::kOwned : 1];
#pragma unroll
            for (int q = 0; q < C::kOwned; ++q) c4[q] = 0;
#pragma unroll
            for (int q = 0; q < (UB ? C::kOwned : 1); ++q) k0[q] = k4[q] = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t* rowp = ring + ((4 * k + r) % kRing) * R::kRowDw;
                // phase ph of band b lies where e ^ (b & 1) == ph: the slots of
                // E for (b ^ ph) odd, of ~E otherwise. `base` is that mask for
                // (b_lo, ph_lo); the other parities use its complement.
                const uint32_t base = window20(rowp, gbit) ^ flip0;
                const uint32_t nbase = base ^ 0xFFFFFu;
#pragma unroll
                for (int hb = 0; hb < C::kBands; ++hb) {
                    const int b = b_lo + hb;
                    if (b >= NB) continue;  // odd NB: the second half is one band short
                    const uint32_t hw = window20(rowp + plane_of<NB, PH>(b, 0) * kPlaneDw, gbit);
                    uint32_t w_first = 0;
#pragma unroll
                    for (int pp = 0; pp < C::kPhases; ++pp) {
                        // the two phases of a band split its hits: the second is
                        // what the first leaves
                        const uint32_t w =
                            pp == 0 ? hw & (((hb ^ pp) & 1) ? nbase : base) : hw ^ w_first;
                        w_first = w;
                        const int q = C::kPhases * hb + pp;
                        c4[q] += __popc(w);
                        if constexpr (UB) {
                            k0[q] += __popc(w & corner_mask(16 + r));
                            k4[q] += __popc(w & corner_mask(r));
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < C::kOwned; ++q) {
                const uint32_t square = (sq[q] & 0xFFFFu) + c4[q] - (sq[q] >> 16);
                sq[q] = square | ((hist[q] >> 24) << 16);
                hist[q] = (hist[q] << 8) | c4[q];
                uint32_t bound = square;
                if constexpr (UB) {
                    bound = square - (k0[q] + (pend[q] & 0xFFu));
                    pend[q] = (pend[q] >> 8) | (k4[q] << 24);
                }
                pass |= (bound >= (uint32_t)p.min_size ? 1u : 0u) << q;
            }
            if (m < 0 || col0 + gbit >= p.phase_cols || s_g >= p.s_end) pass = 0;
            // keep the band-phases with a center inside this job
#pragma unroll
            for (int q = 0; q < C::kOwned; ++q) {
                const int b = b_lo + q / C::kPhases, ph = ph_lo + q % C::kPhases;
                const int64_t ox0 = block_off + (int64_t)ph * kPeriod + col0 + gbit;
                if (b >= NB || ox0 >= p.w || ox0 + 3 < p.ox_min ||
                    (int64_t)b * kPeriod + s_g >= p.h_valid)
                    pass &= ~(1u << q);
            }
        }
        // ---- refinement by the whole wave ----------------------------------
        auto code_of = [&](int q) {
            if constexpr (kMfma) {  // bit 8 s + 4 f + i: D register i of set s, contraction f
                const int md = 4 * (lane >> 4) + (q & 3);
                const int qd = MR::kSets == 2 ? 16 * (q >> 3) + md : md % MR::QN;
                const int hd = MR::kSets == 2 ? 0 : md / MR::QN;
                const int g =
                    8 * (hd * MR::kRangeWords + 4 * wave) + 16 * ((q >> 2) & 1) + (lane & 15);
                return group_code(4 * g, qd >> 1, qd & 1);
            } else if constexpr (kWord)  // bit 8 ph + i: group 8 wd + i, band wb, phase wph + ph
                return group_code(4 * (8 * wd + (q & 7)), wb, wph + (q >> 3));
            else
                return group_code(gbit, b_lo + q / C::kPhases, ph_lo + q % C::kPhases);
        };
        auto emit = [&](uint32_t code, int i, int jr, uint32_t c) {
            const int gb = (int)(code >> 8), b = (int)((code >> 2) & 15u);
            const int ph = (int)(code & (PH == 3 ? 3u : 1u));
            const int64_t ox = block_off + (int64_t)ph * kPeriod + col0 + gb + i;
            const int32_t s_c = s_g + jr;
            if (c < (uint32_t)p.min_size || c > (uint32_t)p.max_size) return;
            if (col0 + gb + i >= p.phase_cols || ox >= p.w || ox < p.ox_min) return;
            if (s_c >= p.s_end || (int64_t)b * kPeriod + s_c >= p.h_valid) return;
            const int32_t cx = (int32_t)((uint32_t)p.x0 + (uint32_t)ox);
            const int32_t cz =
                (int32_t)((uint32_t)p.z0 + (uint32_t)b * (uint32_t)kPeriod + (uint32_t)s_c);
            if (p.use_annulus) {
                const int64_t dx = (int64_t)cx - p.acx, dz = (int64_t)cz - p.acz;
                if (dx * dx + dz * dz < p.rd_min_sq) return;
            }
            atomicAdd(&lhist[c], 1u);
            if (c >= (uint32_t)emit_min) {
                const unsigned long long idx = atomicAdd(p.emitted, 1ull);
                if (idx < p.hit_cap) p.hits[idx] = Hit{(int32_t)c, cx, cz, p.tag};
            }
        };
        // a wave-uniform test first: the refinement's per-lane row addresses are
        // then computed only in the (rare) steps where some lane has a group
        if (__ballot(pass != 0) != 0)
            refine_wave<NB, PH, UB && !kWord>(pass, code_of, ring, 4 * m, p.min_size, emit);
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
