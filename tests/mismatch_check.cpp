// CPU check of the Java-rejection machinery the banded search relies on
// (slime_amd/csrc/mismatch_table.h and row_table_hits in slime_math.h, the
// code the GPU kernel runs) against the oracle:
//   1. the table size, and how many of its states are rejected-then-hit;
//   2. seeds built so that a chosen chunk has a random table state: in a
//      4,096-column window of that row around the chunk, the set of chunks the
//      row search reports equals the oracle's brute-force set of chunks whose
//      first-draw verdict differs from Java's.
#include <cstdint>
#include <cstdio>
#include <random>
#include <set>

#include "mismatch_table.h"

extern "C" int32_t oracle_first_draw(int64_t seed, int64_t cx, int64_t cz);
extern "C" int32_t oracle_is_slime(int64_t seed, int64_t cx, int64_t cz);

using namespace slime;

int main() {
    const std::vector<uint64_t> t = build_mismatch_table();
    long java_hits = 0;
    for (uint64_t s : t) {
        const uint64_t st0 = (s ^ kScramble) & kMask48;
        const uint64_t st1 = (st0 * kJavaMul + 11u) & kMask48;
        java_hits += ((st1 >> 17) % 10u) != 0;  // first draw says no -> Java says yes
    }
    std::mt19937_64 rng(77);
    long windows = 0, equal = 0, planted_found = 0;
    for (int it = 0; it < 400; ++it) {
        const uint64_t e = t[rng() % t.size()];
        const int32_t x = (int32_t)(uint32_t)rng(), z = (int32_t)(uint32_t)rng();
        const int64_t seed =
            (int64_t)(((e - x_term(x) - z_term(z)) & kMask48) | ((rng() & 0xFFFF) << 48));
        int64_t x0 = (int64_t)x - (int64_t)(rng() % 4096);
        if (x0 < INT32_MIN) x0 = INT32_MIN;
        if (x0 + 4096 > (int64_t)INT32_MAX + 1) x0 = (int64_t)INT32_MAX + 1 - 4096;
        int32_t xs[8];
        const int n = row_table_hits(seed, z, t.data(), (int32_t)t.size(), (int32_t)x0, 4096, xs, 8);
        std::set<int32_t> got(xs, xs + (n < 8 ? n : 8)), want;
        for (int64_t cx = x0; cx < x0 + 4096; ++cx)
            if (oracle_first_draw(seed, cx, z) != oracle_is_slime(seed, cx, z)) want.insert((int32_t)cx);
        ++windows;
        equal += got == want && n <= 8;
        planted_found += got.count(x);
    }
    printf("%zu %ld %ld %ld %ld\n", t.size(), java_hits, windows, equal, planted_found);
    return 0;
}
