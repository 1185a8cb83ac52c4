// CPU check of band_tile_of (slime_amd/csrc/search_util.h): the tile of a band
// job that decides a center, which the engines use to apply exact recounts to
// the centers a cancelled search has already decided. For random regions, x
// layouts (one phase; two phases with whole blocks, a right-aligned tail
// block or one masked block; two phases plus a one-phase strip; three phases
// over part of the width), band groups and slot-row cuts:
//   * every center of the region lies in exactly one job, and band_tile_of
//     names a tile whose slot columns and slot rows, decomposed as the kernel
//     does (tile % tiles_x -> x-block and x-tile, tile / tiles_x -> z-tile),
//     hold that center;
//   * centers outside the region lie in no job;
//   * a center drawn inside a random tile's own extent maps back to it.
// Prints: centers checked, tiles checked, failures.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "search_util.h"

using namespace slime;

namespace {

// The kernel's view of tile t of job p (band_tile.h): slot columns
// [col0, col0 + 496) of x-block xblock, slot rows [s0, s0 + tile_h).
struct TileExtent {
    int64_t xblock, col0, s0;
};

// Columns of an x-block: ph P (one block of 2P span with one phase).
int64_t block_cols(int ph) { return (ph == 3 ? 3 : 2) * kPeriod; }

TileExtent extent(const BandParams& p, int64_t t) {
    const int64_t tx = t % p.tiles_x, ts = t / p.tiles_x;
    return TileExtent{tx / p.tiles_per_block, p.slot_lo + (tx % p.tiles_per_block) * kBandTileOut,
                      p.s_begin + ts * p.tile_h};
}

// Whether the kernel's tile t decides center (ox, oz) (offsets from the
// job's x0, z0).
bool tile_holds(const BandParams& p, int nb, int ph, int64_t t, int64_t ox, int64_t oz) {
    const TileExtent e = extent(p, t);
    if (ox < p.ox_min || ox >= p.w) return false;
    bool col = false;
    for (int phase = 0; phase < ph; ++phase) {
        const int64_t c = ox - e.xblock * block_cols(ph) - phase * kPeriod;
        col = col || (c >= e.col0 && c < e.col0 + kBandTileOut && c < p.phase_cols);
    }
    if (!col) return false;
    for (int b = 0; b < nb; ++b) {
        const int64_t s = oz - (int64_t)b * kPeriod;
        if (s >= e.s0 && s < e.s0 + p.tile_h && s < p.s_end && (int64_t)b * kPeriod + s < p.h_valid)
            return true;
    }
    return false;
}

}  // namespace

int main() {
    std::mt19937_64 rng(20261204);
    long centers = 0, tiles = 0, bad = 0;
    for (int it = 0; it < 400; ++it) {
        SearchSpec spec;
        spec.seed = (int64_t)rng();
        const int64_t widths[] = {600,           kPeriod,          kPeriod + 1,
                                  kPeriod + 5000, 2 * kPeriod,      2 * kPeriod + 700,
                                  3 * kPeriod + 1000, 5 * kPeriod - 3, 7 * kPeriod + 5,
                                  6 * kPeriod + 3, 4 * kPeriod,     3 * kPeriod};
        spec.w = widths[rng() % 12];
        spec.h = 1 + (int64_t)(rng() % (4 * (uint64_t)kPeriod));
        spec.x0 = (int32_t)((int64_t)(rng() % 2000000000ull) - 1000000000);
        spec.z0 = (int32_t)((int64_t)(rng() % 2000000000ull) - 1000000000);
        std::vector<XPart> parts;
        const int layout = (int)(rng() % 4);
        if (layout == 2 && spec.w > kPeriod && spec.w % (2 * kPeriod) != 0)
            parts = x_parts_one_phase_tail(spec);
        else if (layout == 3 && spec.w > 2 * kPeriod)
            parts = x_parts(spec, 3);
        else
            parts = x_parts(spec, (layout >= 1 && spec.w > kPeriod) ? 2 : 1);
        const int64_t nbz = bands_of(spec);
        std::vector<int> groups;
        for (int64_t left = nbz; left > 0;) {
            const int g = (int)std::min<int64_t>(left, 1 + (int64_t)(rng() % 4));
            groups.push_back(g);
            left -= g;
        }
        const int32_t tile_h = 4 * (int32_t)(1 + rng() % 64);
        const int64_t max_centers = (int64_t)1 << (30 + rng() % 10);
        std::vector<BandJob> jobs;
        int64_t b_first = 0;
        for (const int nb : groups) {
            for (const XPart& xp : parts)
                band_jobs_part(spec, xp, b_first, nb, slot_rows_of(spec), tile_h, max_centers, jobs);
            b_first += nb;
        }
        for (int k = 0; k < 300; ++k) {
            // a center inside the region (or, one time in six, just outside)
            const bool outside = k % 6 == 5;
            int64_t ox = (int64_t)(rng() % (uint64_t)spec.w), oz = (int64_t)(rng() % (uint64_t)spec.h);
            if (outside) {
                if (rng() & 1) ox = (rng() & 1) ? -1 - (int64_t)(rng() % 9) : spec.w + (int64_t)(rng() % 9);
                else oz = (rng() & 1) ? -1 - (int64_t)(rng() % 9) : spec.h + (int64_t)(rng() % 9);
            }
            const int64_t cx = (int64_t)spec.x0 + ox, cz = (int64_t)spec.z0 + oz;
            int owners = 0;
            for (const BandJob& j : jobs) {
                const int64_t t = band_tile_of(j.p, j.nb, j.ph, cx, cz);
                if (t < 0) continue;
                ++owners;
                if (t >= j.p.tiles_total ||
                    !tile_holds(j.p, j.nb, j.ph, t, cx - j.p.x0, cz - j.p.z0))
                    ++bad;
            }
            if (owners != (outside ? 0 : 1)) ++bad;
            ++centers;
        }
        for (const BandJob& j : jobs) {
            for (int k = 0; k < 8; ++k) {
                const int64_t t = (int64_t)(rng() % (uint64_t)j.p.tiles_total);
                const TileExtent e = extent(j.p, t);
                // a random center of the tile's extent that the job decides
                for (int tries = 0; tries < 64; ++tries) {
                    const int phase = (int)(rng() % (uint64_t)j.ph);
                    const int64_t ox = e.xblock * block_cols(j.ph) + phase * kPeriod + e.col0 +
                                       (int64_t)(rng() % kBandTileOut);
                    const int64_t oz = (int64_t)(rng() % (uint64_t)j.nb) * kPeriod + e.s0 +
                                       (int64_t)(rng() % (uint64_t)j.p.tile_h);
                    if (!tile_holds(j.p, j.nb, j.ph, t, ox, oz)) continue;
                    if (band_tile_of(j.p, j.nb, j.ph, j.p.x0 + ox, j.p.z0 + oz) != t) ++bad;
                    ++tiles;
                    break;
                }
            }
        }
    }
    std::printf("%ld %ld %ld\n", centers, tiles, bad);
    return 0;
}
