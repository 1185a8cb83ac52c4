// Host-side pieces of a search shared by the per-search engine and the
// multi-search sweep: band-job geometry, the first-draw fix list, ranking.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "engine.h"
#include "scan.h"
#include "slime_math.h"

namespace slime {

// Largest score s with at least k counted hits of score >= s; `floor` if none.
inline int32_t gate_for(const int64_t* hist, int64_t k, int32_t floor) {
    int64_t acc = 0;
    for (int32_t s = kHistBins - 1; s >= 0; --s) {
        acc += std::max<int64_t>(0, hist[s]);
        if (acc >= k) return std::max(s, floor);
    }
    return floor;
}

// The first-draw mismatch chunks of one search, sorted by (z, x).
struct MismatchIndex {
    std::vector<MismatchChunk> mm;

    static bool in_disk(int64_t dx, int64_t dz) { return dx * dx + dz * dz <= 68; }

    void sort() {
        std::sort(mm.begin(), mm.end(), [](const MismatchChunk& a, const MismatchChunk& b) {
            return a.z != b.z ? a.z < b.z : a.x < b.x;
        });
    }

    // Does the disk of center (cx, cz) contain a mismatch chunk?
    bool affects(int32_t cx, int32_t cz) const {
        if (mm.empty()) return false;
        auto it = std::lower_bound(mm.begin(), mm.end(), (int64_t)cz - 8,
                                   [](const MismatchChunk& m, int64_t z) { return m.z < z; });
        for (; it != mm.end() && (int64_t)it->z <= (int64_t)cz + 8; ++it)
            if (in_disk((int64_t)it->x - cx, (int64_t)it->z - cz)) return true;
        return false;
    }
};

// Every center of the spec whose disk holds a mismatch chunk (index sorted),
// once each, appended to `fixes` (a container with reserve and push_back) as
// fixes to recount.
template <typename Fixes>
inline void collect_fixes(const SearchSpec& spec, const MismatchIndex& index, Fixes& fixes) {
    const std::vector<MismatchChunk>& mm = index.mm;
    const int64_t rmin_sq = spec.rd_min * spec.rd_min;
    fixes.reserve(fixes.size() + mm.size() * (size_t)kDiskCells);
    for (size_t i = 0; i < mm.size(); ++i) {
        const MismatchChunk& m = mm[i];
        // earlier mismatches close enough to share a center already added it
        size_t first_near = i;
        while (first_near > 0 && (int64_t)mm[first_near - 1].z >= (int64_t)m.z - 16) --first_near;
        for (int dz = -8; dz <= 8; ++dz) {
            const int hw = disk_half_width(dz < 0 ? -dz : dz);
            for (int dx = -hw; dx <= hw; ++dx) {
                const int64_t cx = (int64_t)m.x + dx, cz = (int64_t)m.z + dz;
                const int64_t ox = cx - spec.x0, oz = cz - spec.z0;
                if (ox < 0 || ox >= spec.w || oz < 0 || oz >= spec.h) continue;
                if (spec.band_rows > 0 && oz % kPeriod >= spec.band_rows) continue;
                if (spec.rd_min > 0) {
                    const int64_t ax = cx - spec.acx, az = cz - spec.acz;
                    if (ax * ax + az * az < rmin_sq) continue;
                }
                bool seen = false;
                for (size_t j = first_near; j < i && !seen; ++j)
                    seen = MismatchIndex::in_disk((int64_t)mm[j].x - cx, (int64_t)mm[j].z - cz);
                if (!seen)
                    fixes.push_back(CenterFix{spec.seed, (int32_t)cx, (int32_t)cz, 0, 0});
            }
        }
    }
}

// Ranks the candidates (size desc, x asc, z asc) and writes the best `keep`
// as reference rows (block coordinates cx*16+8, precise_afk == 0 columns).
inline size_t write_rows(std::vector<Hit>& cand, size_t keep, ResultRow* out) {
    const size_t n = std::min(cand.size(), keep);
    if (n < cand.size())
        std::partial_sort(cand.begin(), cand.begin() + n, cand.end(), ranks_before);
    else
        std::sort(cand.begin(), cand.end(), ranks_before);
    for (size_t i = 0; i < n; ++i) {
        const int64_t bx = (int64_t)cand[i].cx * 16 + 8;
        const int64_t bz = (int64_t)cand[i].cz * 16 + 8;
        out[i].size = cand[i].size;
        out[i].center_x = (int32_t)(uint32_t)(uint64_t)bx;
        out[i].center_z = (int32_t)(uint32_t)(uint64_t)bz;
        out[i].obs_count = 0;
        out[i].afk_x = out[i].center_x;
        out[i].afk_z = out[i].center_z;
    }
    return n;
}

inline void fill_filters(const SearchSpec& spec, BandParams& p) {
    p.min_size = spec.min_size;
    p.max_size = spec.max_size;
    p.emit_min = spec.min_size;
    p.use_annulus = spec.rd_min > 0;
    p.rd_min_sq = spec.rd_min * spec.rd_min;
    p.acx = spec.acx;
    p.acz = spec.acz;
}

// ---- band-job geometry -------------------------------------------------------
//
// Along x, two phases (ph = 2, only for regions wider than P): whole 2P column
// blocks from x0 form the main part; a remainder becomes one block aligned to
// the right edge whose tiles start at the first slot column that reaches a new
// center, with the already covered centers masked out; a region between P and
// 2P wide is one block whose phase-1 columns past the edge are masked out.
// One phase (ph = 1): a single part spanning the region.
// Three phases (ph = 3, regions wider than 2P): phases x, x + P, x + 2P over
// 3P blocks; see x_parts_three.
struct XPart {
    int32_t x0, w, ox_min, slot_lo, nblocks, phase_cols, ph;
    int64_t centers_per_row;

    int32_t tiles_per_block() const {
        return (phase_cols - slot_lo + kBandTileOut - 1) / kBandTileOut;
    }
};

inline std::vector<XPart> x_parts_three(const SearchSpec& spec);

inline std::vector<XPart> x_parts(const SearchSpec& spec, int ph) {
    const int64_t W = spec.w;
    if (ph == 3) return x_parts_three(spec);
    std::vector<XPart> parts;
    if (ph == 2) {
        const int64_t nfull = W / (2 * kPeriod), rem = W % (2 * kPeriod);
        if (nfull == 0) {
            parts.push_back(XPart{spec.x0, (int32_t)W, 0, 0, 1, kPeriod, 2, W});
        } else {
            parts.push_back(XPart{spec.x0, (int32_t)(nfull * 2 * kPeriod), 0, 0, (int32_t)nfull,
                                  kPeriod, 2, nfull * 2 * kPeriod});
            if (rem > 0)
                parts.push_back(XPart{(int32_t)((uint32_t)spec.x0 + (uint32_t)(W - 2 * kPeriod)),
                                      2 * kPeriod, (int32_t)(2 * kPeriod - rem),
                                      (int32_t)std::max<int64_t>(0, kPeriod - rem), 1, kPeriod, 2,
                                      rem});
        }
    } else {
        parts.push_back(XPart{spec.x0, (int32_t)W, 0, 0, 1, (int32_t)W, 1, W});
    }
    return parts;
}

// Three phases for a region wider than 2P, every column decided once: with
// W = 2P n + r, the last 2P + r columns (0 < r <= P) are three phases over
// slot columns [0, r) and two over [r, P); for P < r < 2P a full 3P block
// takes P + r of them first (the rest as above, or with n = 1 a one-phase
// strip of r - P); whole 2P blocks before them keep two phases. With r = 0
// the two-phase layout is already exact.
inline std::vector<XPart> x_parts_three(const SearchSpec& spec) {
    const int64_t W = spec.w, P = kPeriod;
    const int64_t n = W / (2 * P), r = W % (2 * P);
    std::vector<XPart> parts;
    if (n == 0 || r == 0) return x_parts(spec, W > P ? 2 : 1);
    auto at = [&](int64_t off) { return (int32_t)((uint32_t)spec.x0 + (uint32_t)off); };
    const int64_t n2 = r <= P ? n - 1 : n - 2;  // whole two-phase blocks first
    if (n2 > 0)
        parts.push_back(XPart{spec.x0, (int32_t)(n2 * 2 * P), 0, 0, (int32_t)n2, (int32_t)P, 2,
                              n2 * 2 * P});
    int64_t off = std::max<int64_t>(0, n2) * 2 * P;
    int64_t rr = r;  // columns past a 2P block still to cover with three phases
    if (r > P) {
        parts.push_back(XPart{at(off), (int32_t)(3 * P), 0, 0, 1, (int32_t)P, 3, 3 * P});
        off += 3 * P;
        rr = r - P;
        if (n == 1) {  // no 2P block left to pair with: one phase over the rest
            parts.push_back(XPart{at(off), (int32_t)rr, 0, 0, 1, (int32_t)rr, 1, rr});
            return parts;
        }
    }
    // the last 2P + rr columns: three phases over slot columns [0, rr), two
    // over [rr, P)
    parts.push_back(XPart{at(off), (int32_t)(2 * P + rr), 0, 0, 1, (int32_t)rr, 3, 3 * rr});
    if (rr < P)
        parts.push_back(XPart{at(off), (int32_t)(2 * P), 0, (int32_t)rr, 1, (int32_t)P, 2,
                              2 * (P - rr)});
    return parts;
}

// Slot rows per band: min(h, P), or the band_rows of a strided band set.
inline int32_t slot_rows_of(const SearchSpec& spec) {
    int64_t r = std::min<int64_t>(spec.h, kPeriod);
    if (spec.band_rows > 0) r = std::min<int64_t>(r, spec.band_rows);
    return (int32_t)r;
}

inline int64_t bands_of(const SearchSpec& spec) { return (spec.h + kPeriod - 1) / kPeriod; }

struct BandJob {
    BandParams p;
    int nb, ph;
    int64_t centers;  // centers this job decides
};

// Jobs of one x-part and one group of nb consecutive bands from band b_first,
// each over slot rows [s_lo, s_rows), cut into slot-row ranges of about
// max_centers centers. Output pointers are left null.
inline void band_jobs_part(const SearchSpec& spec, const XPart& xp, int64_t b_first, int nb,
                           int32_t s_rows, int32_t tile_h, int64_t max_centers,
                           std::vector<BandJob>& out, int32_t s_lo = 0) {
    const int32_t tpb = xp.tiles_per_block();
    int64_t rows = max_centers / std::max<int64_t>(1, nb * xp.centers_per_row);
    rows = std::max<int64_t>(tile_h, rows / tile_h * tile_h);
    for (int64_t s = s_lo; s < s_rows; s += rows) {
        BandJob j{};
        j.nb = nb;
        j.ph = xp.ph;
        BandParams& p = j.p;
        p.seed = spec.seed;
        p.x0 = xp.x0;
        p.w = xp.w;
        p.ox_min = xp.ox_min;
        p.slot_lo = xp.slot_lo;
        p.phase_cols = xp.phase_cols;
        p.tiles_per_block = tpb;
        p.z0 = (int32_t)((uint32_t)spec.z0 + (uint32_t)(b_first * kPeriod));
        p.h_valid = (int32_t)std::min<int64_t>(spec.h - b_first * kPeriod, (int64_t)nb * kPeriod);
        p.s_begin = (int32_t)s;
        p.s_end = (int32_t)std::min<int64_t>(s_rows, s + rows);
        p.tile_h = tile_h;
        p.tiles_x = xp.nblocks * tpb;
        p.tiles_total = p.tiles_x * ((p.s_end - p.s_begin + tile_h - 1) / tile_h);
        fill_filters(spec, p);
        int64_t c = 0;
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t hi = std::min<int64_t>(p.s_end, p.h_valid - b * kPeriod);
            if (hi > p.s_begin) c += (hi - p.s_begin) * xp.centers_per_row;
        }
        j.centers = c;
        out.push_back(j);
    }
}

// Jobs over every center of the spec: for each group of consecutive bands
// (sizes in `groups`, summing to bands_of(spec)) and each x-part, all slot
// rows of the spec.
inline void band_jobs(const SearchSpec& spec, const std::vector<XPart>& parts,
                      const std::vector<int>& groups, int32_t tile_h, int64_t max_centers,
                      std::vector<BandJob>& out) {
    const int32_t slot_rows = slot_rows_of(spec);
    int64_t b_first = 0;
    for (const int nb : groups) {
        for (const XPart& xp : parts)
            band_jobs_part(spec, xp, b_first, nb, slot_rows, tile_h, max_centers, out);
        b_first += nb;
    }
}

// The tile of a band job that decides center (cx, cz), as an index into the
// job's tiles (what band_tile's `tile` argument and the sweep's done counts
// refer to), or -1 when the job does not cover that center. Tile t holds
// x-tile tx = t % tiles_x (x-block tx / tiles_per_block, 496 slot columns
// from slot_lo + (tx % tiles_per_block) * 496, every phase) and slot rows
// s_begin + (t / tiles_x) * tile_h onwards, in every band of the job.
inline int64_t band_tile_of(const BandParams& p, int nb, int ph, int64_t cx, int64_t cz) {
    const int64_t ox = cx - (int64_t)p.x0;
    if (ox < p.ox_min || ox >= p.w) return -1;
    int64_t xblock = 0, c = ox;
    if (ph >= 2) {  // x-blocks of ph P columns, the slot column is ox mod P
        xblock = ox / (ph * (int64_t)kPeriod);
        c = (ox - xblock * ph * kPeriod) % kPeriod;
    }
    if (c < p.slot_lo || c >= p.phase_cols) return -1;
    const int64_t oz = cz - (int64_t)p.z0;
    if (oz < 0 || oz >= p.h_valid) return -1;
    const int64_t b = oz / kPeriod, s = oz - b * kPeriod;
    if (b >= nb || s < p.s_begin || s >= p.s_end) return -1;
    const int64_t xtile = (c - p.slot_lo) / kBandTileOut;
    const int64_t ts = (s - p.s_begin) / p.tile_h;
    return ts * p.tiles_x + xblock * p.tiles_per_block + xtile;
}

// Whether a plain scan launch decides center (cx, cz).
inline bool plain_covers(const ScanParams& p, int64_t cx, int64_t cz) {
    const int64_t ox = cx - (int64_t)p.x0, oz = cz - (int64_t)p.z0;
    return ox >= 0 && ox < p.w && oz >= 0 && oz < p.h;
}

// The largest region one search may cover, in centers. Its first-draw
// mismatch list (about 209,712 chunks per 2^48) must fit kMismatchCap and its
// recount list (up to 221 centers per mismatch, 24 bytes each: about 1.1 GB
// at the limit) host memory. 2^48 centers is the square of radius 8,388,607
// chunks, 4.5 times the world border's 1,875,000; the reference accepts radii
// up to (2^31 - 20) / 2, which at its published rate would scan for weeks.
constexpr int64_t kMaxSearchCenters = int64_t(1) << 48;
inline bool within_search_limit(int64_t w, int64_t h) {
    return w > 0 && h > 0 && w <= kMaxSearchCenters / h;
}

// Two phases where both phases of a slot column lie inside the region, one
// phase (its bands paired) elsewhere: whole 2P blocks from x0; then, for the
// remainder r = W mod 2P, a one-phase strip of its r columns when r <= P, or
// else two phases over its slot columns [0, r - P) and one phase over slot
// columns [r - P, P). For regions wider than P that are not a whole number of
// 2P blocks wide.
inline std::vector<XPart> x_parts_one_phase_tail(const SearchSpec& spec) {
    const int64_t W = spec.w, P = kPeriod;
    const int64_t nfull = W / (2 * P), rem = W % (2 * P);
    std::vector<XPart> parts;
    if (nfull > 0)
        parts.push_back(XPart{spec.x0, (int32_t)(nfull * 2 * P), 0, 0, (int32_t)nfull, (int32_t)P,
                              2, nfull * 2 * P});
    if (rem == 0) return parts;
    const int32_t at = (int32_t)((uint32_t)spec.x0 + (uint32_t)(nfull * 2 * P));
    if (rem <= P) {
        parts.push_back(XPart{at, (int32_t)rem, 0, 0, 1, (int32_t)rem, 1, rem});
    } else {
        const int64_t r2 = rem - P;  // slot columns with both phases inside
        parts.push_back(XPart{at, (int32_t)rem, 0, 0, 1, (int32_t)r2, 2, 2 * r2});
        parts.push_back(XPart{(int32_t)((uint32_t)at + (uint32_t)r2), (int32_t)(P - r2), 0, 0, 1,
                              (int32_t)(P - r2), 1, P - r2});
    }
    return parts;
}

}  // namespace slime
