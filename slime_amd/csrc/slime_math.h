// Integer arithmetic of the slime-chunk test, shared by host and device code.
//
// Chunk (x, z) of world seed S is a slime chunk when Java's
//   new Random(S + xt(x) + zt(z) ^ 0x3AD8025F).nextInt(10) == 0
// (SlimeCore.cpp:208-232). Only the low 48 bits of the sum reach the LCG.
//
// The first LCG output is split as Y = (state1 >> 16) mod 2^32, so that
// next(31) == Y >> 1 and the first-draw test is (Y >> 1) % 10 == 0. Java
// redraws when next(31) >= 2^31 - 8, i.e. when Y >= 0xFFFFFFF0; those states
// take the full redraw loop (SlimeCore.cpp:54-56, 74-79).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SLIME_HD __host__ __device__ __forceinline__
#else
#define SLIME_HD static inline
#endif

namespace slime {

constexpr uint64_t kMask48 = 0xFFFFFFFFFFFFull;
constexpr uint64_t kJavaMul = 0x5DEECE66Dull;      // java.util.Random multiplier
constexpr uint32_t kJavaMulLo = 0xDEECE66Du;       // its low 32 bits
constexpr uint64_t kScramble = 0x5DEECE66Dull ^ 0x3AD8025Full;  // = 0x5E434E432
constexpr uint32_t kRejectY = 0xFFFFFFF0u;         // Y >= this: next(31) >= 2^31-8
constexpr uint32_t kInv5 = 0xCCCCCCCDu;            // 5^-1 mod 2^32

// Half-widths of the 221-cell disk rows, dz = -8..8 (dx^2 + dz^2 <= 68).
constexpr int kDiskRows = 17;
constexpr int kDiskCells = 221;

SLIME_HD int disk_half_width(int adz) {
    // 8,8,8,7,7,6,5,4,2 for |dz| = 0..8
    return adz <= 2 ? 8 : adz <= 4 ? 7 : adz == 5 ? 6 : adz == 6 ? 5 : adz == 7 ? 4 : 2;
}

// xt(x) = (int)(x*x*4987142) + (int)(x*5947611), as a 64-bit two's complement.
SLIME_HD uint64_t x_term(int32_t x) {
    uint32_t u = (uint32_t)x;
    int32_t a = (int32_t)(u * u * 4987142u);
    int32_t b = (int32_t)(u * 5947611u);
    return (uint64_t)((int64_t)a + (int64_t)b);
}

// zt(z) = (long)(int)(z*z) * 4392871 + (int)(z*389711).
SLIME_HD uint64_t z_term(int32_t z) {
    uint32_t u = (uint32_t)z;
    int64_t sq = (int64_t)(int32_t)(u * u);
    int32_t lin = (int32_t)(u * 389711u);
    return (uint64_t)(sq * 4392871ll) + (uint64_t)(int64_t)lin;
}

// Y = ((s ^ scramble) * 0x5DEECE66D + 11) >> 16, mod 2^32, from s = seed + xt + zt.
// Written on 32-bit halves: with state0 = hi*2^32 + lo (hi < 2^16 matters),
//   state1 >> 16 = ((lo*Mlo + 11) >> 16) + ((lo*5 + hi*Mlo) << 16)  (mod 2^32).
SLIME_HD uint32_t first_y(uint64_t s) {
    uint64_t st0 = s ^ kScramble;
    uint32_t lo = (uint32_t)st0;
    uint32_t hi = (uint32_t)(st0 >> 32);
    uint64_t p = (uint64_t)lo * kJavaMulLo + 11u;
    uint32_t low_part = (uint32_t)(p >> 16);
    uint32_t cross = lo * 5u + hi * kJavaMulLo;
    return low_part + (cross << 16);
}

// First-draw test: next(31) % 10 == 0, ignoring Java's rejection.
SLIME_HD bool first_draw_hit(uint32_t y) {
    uint32_t bits = y >> 1;
    return bits % 10u == 0u;
}

// Java's full nextInt(10) == 0 from the first LCG state, for the rare
// rejected first draws (Y >= kRejectY).
SLIME_HD bool redraw_hit(uint64_t s) {
    uint64_t st = ((s ^ kScramble) * kJavaMul + 11u) & kMask48;
    for (;;) {
        uint32_t bits = (uint32_t)(st >> 17);
        uint32_t val = bits % 10u;
        if (bits - val + 9u <= 0x7FFFFFFFu) return val == 0u;
        st = (st * kJavaMul + 11u) & kMask48;
    }
}

// ---- slot arithmetic of the banded scan ----------------------------------
//
// A slot is a phase-0 column x with X0 = seed + xt(x), X1 = seed + xt(x+2^17),
// and a row whose z-terms for the bands b = 0.. are Z_b = zt(z + b*2^17). All
// 2 x NB seed sums agree mod 2^16 (also in the carry out of the low half) and
// their first-draw parity flips with the phase and with the band, so one test
// per band decides both phases: the phase p = e ^ (b & 1) is the only one that
// can be even.
constexpr uint32_t kLowScramble = 0xE4320000u;     // (0x5E434E432 & 0xFFFF) << 16
constexpr uint32_t kHighScramble = 0x5E434u;       // 0x5E434E432 >> 16
constexpr uint32_t kMulInv5 = kJavaMulLo * kInv5;  // M * 5^-1 (mod 2^32)
constexpr uint32_t kFifthLimit = 858993459u;       // (2^32 - 1) / 5

struct BandSlot {
    uint32_t xa;     // high half of the phase even in even bands (no carry)
    uint32_t xb;     // the other phase
    uint32_t carry;  // carry out of the low half, common to all chunks of the slot
    uint32_t bias;   // u + (w - (Y & 1)) * 5^-1
    bool e;          // first-draw parity of phase 0, band 0
};

// xlo = low 16 bits of X0 in the top half; xh0/xh1 = bits 16..47 of X0/X1;
// zlo, zhi0 likewise of Z_0.
SLIME_HD BandSlot band_slot(uint32_t xlo, uint32_t xh0, uint32_t xh1, uint32_t zlo,
                            uint32_t zhi0) {
    const uint64_t lsum = (uint64_t)xlo + zlo;
    const uint32_t carry = (uint32_t)(lsum >> 32);
    const uint32_t u = (uint32_t)lsum ^ kLowScramble;  // low 16 state bits << 16
    // w = ((low16 * 0xDEECE66D + 11) >> 16), the low half's share of Y
    const uint32_t w = (uint32_t)(((uint64_t)u * kJavaMulLo + (11u << 16)) >> 32);
    const uint32_t t = xh0 + carry + zhi0 + w;  // Y (phase 0, band 0) mod 4 == t mod 4
    BandSlot s;
    s.e = (t & 2u) != 0;
    // the phases are selected before the carry is added: the carry joins each
    // band's sum (one three-operand add) instead of costing an add per phase
    s.xa = s.e ? xh1 : xh0;
    s.xb = s.e ? xh0 : xh1;
    s.carry = carry;
    s.bias = u + (w - (t & 1u)) * kInv5;
    return s;
}

// (Y - (Y & 1)) * 5^-1 of band b's tested chunk; next(31) % 10 == 0 iff the
// result is <= kFifthLimit.
SLIME_HD uint32_t band_value(const BandSlot& s, int b, uint32_t zhi_b) {
    const uint32_t a = (((b & 1) ? s.xb : s.xa) + zhi_b + s.carry) ^ kHighScramble;
    return a * kMulInv5 + s.bias;
}

// ---- Java-rejected first draws --------------------------------------------
//
// The x (mod 2^32) with x*x*4987142 + x*5947611 == v (mod 2^32): the map is a
// permutation of Z/2^32 (its derivative 2*4987142*x + 5947611 is odd), so bit
// k of x is fixed by bit k of the residual once bits 0..k-1 are.
SLIME_HD uint32_t invert_x_low32(uint32_t v) {
    uint32_t x = 0;
    for (int k = 0; k < 32; ++k) {
        const uint32_t f = x * x * 4987142u + x * 5947611u;
        x |= ((f ^ v) >> k & 1u) << k;
    }
    return x;
}

// Chunks of row z with x in [x_first, x_first + cols) whose seed sum is in the
// sorted table (n entries, values mod 2^48). xt(x) lies in [-2^32, 2^32), so
// only entries within 2^33 above seed + zt(z) - 2^32 can occur; each names one
// int32 x. Writes up to max_out xs; returns how many it found.
SLIME_HD int row_table_hits(int64_t seed, int32_t z, const uint64_t* table, int32_t n,
                            int32_t x_first, int64_t cols, int32_t* xs, int max_out) {
    const uint64_t base = ((uint64_t)seed + z_term(z)) & kMask48;
    const uint64_t lo = (base - (1ull << 32)) & kMask48;
    int32_t k = 0, hi = n;  // first entry >= lo
    while (k < hi) {
        const int32_t mid = (k + hi) >> 1;
        if (table[mid] < lo) k = mid + 1; else hi = mid;
    }
    int found = 0;
    for (int32_t c = 0; c < n; ++c, ++k) {
        if (k >= n) k -= n;  // the window may wrap past 2^48
        const uint64_t off = (table[k] - lo) & kMask48;
        if (off >= (1ull << 33)) break;
        const int64_t d = (int64_t)off - (1ll << 32);  // required xt(x)
        const int32_t x = (int32_t)invert_x_low32((uint32_t)(uint64_t)d);
        const int64_t dx = (int64_t)x - (int64_t)x_first;
        if (dx < 0 || dx >= cols || (int64_t)x_term(x) != d) continue;
        if (found < max_out) xs[found] = x;
        ++found;
    }
    return found;
}

SLIME_HD bool is_slime(int64_t seed, int32_t x, int32_t z) {
    uint64_t s = (uint64_t)seed + x_term(x) + z_term(z);
    uint32_t y = first_y(s);
    if (y >= kRejectY) return redraw_hit(s);
    return first_draw_hit(y);
}

}  // namespace slime
