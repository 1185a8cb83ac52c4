// Host-side table of the seed sums whose first draw disagrees with Java.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "slime_math.h"

namespace slime {

// Inverse of the Java LCG multiplier mod 2^64 (Newton iteration, odd input).
constexpr uint64_t inverse_mod_2_64(uint64_t a) {
    uint64_t x = a;  // correct to 3 bits for odd a
    for (int i = 0; i < 6; ++i) x *= 2 - a * x;
    return x;
}
constexpr uint64_t kJavaMulInv = inverse_mod_2_64(kJavaMul);
static_assert(((kJavaMul * kJavaMulInv) & kMask48) == 1, "LCG inverse");

// Java rejects a first draw when next(31) >= 2^31 - 8, i.e. for the top 2^20
// first-LCG states (SlimeCore.cpp:54-56). Each is followed through the redraw
// loop; where Java's verdict differs from "next(31) % 10 == 0" the state is
// mapped back through the LCG step and the seed scramble to the seed sum
// s = seed + xt(x) + zt(z) (mod 2^48). Sorted; 209,712 entries.
inline std::vector<uint64_t> build_mismatch_table() {
    std::vector<uint64_t> t;
    const uint64_t first = ((uint64_t)(0x7FFFFFFFu - 7u)) << 17;
    for (uint64_t st1 = first; st1 <= kMask48; ++st1) {
        uint32_t bits = (uint32_t)(st1 >> 17);
        const bool first_draw = bits % 10u == 0u;
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

}  // namespace slime
