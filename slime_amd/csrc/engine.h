// Host-side search engine behind the C ABI.
#pragma once

#include <stdint.h>

#include <atomic>
#include <vector>

#include "scan.h"

namespace slime {

// Result row of the reference ABI (SlimeCore.cpp:182-189), 24 bytes.
struct ResultRow {
    int32_t size;
    int32_t center_x;  // block coordinates cx*16+8
    int32_t center_z;
    int32_t obs_count;
    int32_t afk_x;
    int32_t afk_z;
};
static_assert(sizeof(ResultRow) == 24, "ABI row is 24 bytes");

// One search: a square of radius rd_max around (center_x, center_z), or any
// rectangle of centers, with the reference's per-center filters.
enum class ScanMode : int32_t {
    kAuto = 0,          // banded scan (two x-phases above 2^17 columns, one below)
    kPlain = 1,         // independent per-chunk scan everywhere
    kBanded = 2,        // banded scan (same as kAuto)
    kBandedSquare = 3,  // banded scan, square bound before refinement
    kBandedUnion = 4,   // banded scan, union bound (corner cells off the square)
    kBandedUnion3 = 5,  // as kBandedUnion, three x-phases wherever wider than 2^18
};

// The bound the banded scan uses when the mode does not force one: the union
// bound at every threshold. Its kernels of more than 3 band-phases subtract
// the corner blocks of the square's first and last two rows while counting by
// words, which costs little and leaves fewer groups to refine; measured on the
// full world it is ahead of the square bound from threshold 50 (4.5 vs 2.2 T/s)
// to 62 (8.9 vs 8.6 T/s).
inline bool union_bound_for(ScanMode mode, int32_t /*min_size*/) {
    return mode != ScanMode::kBandedSquare;
}

struct SearchSpec {
    int64_t seed = 0;
    int32_t x0 = 0, z0 = 0;  // first center
    int64_t w = 0, h = 0;    // centers
    int32_t min_size = 0, max_size = 221;
    int64_t rd_min = 0;
    int32_t acx = 0, acz = 0;  // annulus center (search center)
    // 0: the rows [z0, z0+h). Otherwise only the rows z0 + b*2^17 + s with
    // s < band_rows (and b*2^17 + s < h): a strided set of bands, used to
    // exercise the banded scan on small inputs.
    int32_t band_rows = 0;
    ScanMode mode = ScanMode::kAuto;
};

// Process-global control flags of the reference ABI (SlimeCore.cpp:16-27,
// SlimeCoreGPU.cu:3085-3116).
struct Control {
    std::atomic<int> cancel{0};
    std::atomic<int> pause{0};
    std::atomic<int> progress{0};
    std::atomic<int64_t> processed_centers{0};
    std::atomic<int64_t> scan_work_ns{0};
};
Control& control();

// Runs one search to completion on the current device. Returns the total
// number of in-range hits and writes up to max_results ranked rows into out
// (block coordinates, precise_afk == 0 columns); hist_out (optional) gets the
// per-score counts. Returns -1 on a device error.
int64_t run_search(const SearchSpec& spec, ResultRow* out, int32_t max_results,
                   int64_t* hist_out = nullptr);

// Bands per banded launch (1..16); a tuning knob.
void set_max_bands(int nb);

// Device time of the scan kernel launches (HIP events on the engine's stream)
// of one kind (0 plain, 1 banded, 2 both) and the centers they decided.
void kernel_stats(int kind, int64_t* launches, int64_t* kernel_ns, int64_t* centers, bool reset);

// Releases every device buffer the engine holds.
void release_engine();

// The scan kernel that decided most centers of the last search: 16 the plain
// per-chunk kernel, 17 the banded kernel with one x-phase, 18 with two; -1
// before any search.
int last_scan_kernel();

// Ranking order of the reference: size desc, then center_x asc, center_z asc
// compared on 64-bit block coordinates (SlimeCore.cpp:191-206).
inline bool ranks_before(const Hit& a, const Hit& b) {
    if (a.size != b.size) return a.size > b.size;
    if (a.cx != b.cx) return a.cx < b.cx;
    return a.cz < b.cz;
}

}  // namespace slime
