// Device-facing types and launchers of the cluster scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slime {

// One emitted center: exact 221-cell count and chunk coordinates.
struct Hit {
    int32_t size;
    int32_t cx;
    int32_t cz;
};

constexpr int kHistBins = 222;  // exact counts 0..221

// A rectangle of centers [x0, x0+w) x [z0, z0+h) of one seed, plus the
// filters the reference applies per center (SlimeCore.cpp:504-545).
struct ScanParams {
    int64_t seed;
    int32_t x0, z0;       // first center (chunk coordinates)
    int32_t w, h;         // centers
    int32_t tile_h;       // centers per tile along z, multiple of 4
    int32_t tiles_x;      // tiles along x
    int32_t tiles_total;  // tiles in this launch
    int32_t min_size, max_size;
    int32_t emit_min;     // hits below this are counted, not stored
    int32_t use_annulus;  // rd_min > 0
    int64_t rd_min_sq;
    int32_t acx, acz;     // annulus center
    Hit* hits;
    uint32_t hit_cap;
    unsigned long long* emitted;  // number of hits offered for storage
    unsigned long long* hist;     // kHistBins counters of in-range hits
};

// Plain per-chunk scan: tile = 1008 centers x tile_h centers.
constexpr int kTileOut = 1008;
hipError_t launch_scan_plain(const ScanParams& p, hipStream_t stream);

}  // namespace slime
