// CPU check of the banded scan's slot arithmetic (slime_amd/csrc/slime_math.h,
// the functions the kernel calls) against the oracle's first-draw predicate:
// for random seeds and int32 coordinates, every band's single test must give
// the first draw of the selected phase and rule out the other phase.
#include <cstdint>
#include <cstdio>
#include <random>

#include "slime_math.h"

extern "C" int32_t oracle_first_draw(int64_t seed, int64_t cx, int64_t cz);

int main(int argc, char** argv) {
    using namespace slime;
    const long iters = argc > 1 ? atol(argv[1]) : 200000;
    const int nb = 16;
    const uint32_t P = 1u << 17;
    std::mt19937_64 rng(20261203);
    long bad = 0, chunks = 0, hits = 0;
    for (long it = 0; it < iters; ++it) {
        const int64_t seed = (int64_t)rng();
        const int32_t x = (int32_t)(uint32_t)rng(), z = (int32_t)(uint32_t)rng();
        const uint64_t X0 = (uint64_t)seed + x_term(x);
        const uint64_t X1 = (uint64_t)seed + x_term((int32_t)((uint32_t)x + P));
        uint32_t zhi[nb];
        uint32_t zlo = 0;
        for (int b = 0; b < nb; ++b) {
            const uint64_t Z = z_term((int32_t)((uint32_t)z + (uint32_t)b * P));
            zhi[b] = (uint32_t)(Z >> 16);
            if (b == 0) zlo = (uint32_t)Z << 16;
        }
        const BandSlot s = band_slot((uint32_t)X0 << 16, (uint32_t)(X0 >> 16),
                                     (uint32_t)(X1 >> 16), zlo, zhi[0]);
        for (int b = 0; b < nb; ++b) {
            const bool hit = band_value(s, b, zhi[b]) <= kFifthLimit;
            const int sel = (s.e ? 1 : 0) ^ (b & 1);
            for (int p = 0; p < 2; ++p) {
                const int64_t xx = (int32_t)((uint32_t)x + (uint32_t)p * P);
                const int64_t zz = (int32_t)((uint32_t)z + (uint32_t)b * P);
                const bool want = oracle_first_draw(seed, xx, zz) != 0;
                const bool got = p == sel ? hit : false;
                bad += want != got;
                hits += want;
                ++chunks;
            }
        }
    }
    printf("%ld %ld %ld\n", chunks, hits, bad);
    return bad != 0;
}
