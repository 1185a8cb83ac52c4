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

SLIME_HD bool is_slime(int64_t seed, int32_t x, int32_t z) {
    uint64_t s = (uint64_t)seed + x_term(x) + z_term(z);
    uint32_t y = first_y(s);
    if (y >= kRejectY) return redraw_hit(s);
    return first_draw_hit(y);
}

}  // namespace slime
