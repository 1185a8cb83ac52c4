// Precise AFK refinement of ranked result rows on the GPU.
//
// For each row the refinement looks at AFK points around the cluster center
// (81 points, dx and dz in -16..16 blocks step 4, or the center alone) and at
// a list of AFK heights, and counts the spawnable block columns of slime
// chunks: blocks (x, z) with 576 < (x-ox)^2 + (z-oz)^2 + dy^2 <= 16384,
// dy = platform_y - y (SlimeCore.cpp:148-180, 234-331). The best point and
// height are written into obs_count / afk_x / afk_z (SlimeCore.cpp:578-629,
// 654-773; SlimeCoreGPU.cu:2749-2999).
#pragma once

#include <stdint.h>

#include "engine.h"

namespace slime {

// Which reference loop decides ties and the all-zero case.
enum RefineRule : int32_t {
    // Best starts at (0, center, platform_y); only a strictly larger count
    // replaces it (SlimeCore.cpp:582-606, refine_afk_kernel
    // SlimeCoreGPU.cu:2749-2784).
    kRuleKeepCenter = 0,
    // Best starts at -1; equal counts go to the lowest (offset, level) order,
    // so an all-zero row takes offset 0 and level 0 (refine_afk_block_kernel
    // and refine_afk_xz_kernel, SlimeCoreGPU.cu:2799-2861, 2906-2968).
    kRuleFirstOrder = 1,
};

// The AFK heights: level i is y_start + i * y_step, i < y_count.
struct YLevels {
    int32_t platform_y = -64;
    int64_t y_start = -64;
    int64_t y_step = 1;
    int64_t y_count = 1;
};

// The level list of a y-scan configuration (SlimeCore.cpp:148-159,
// SlimeCoreGPU.cu:78-86): y_min..y_max step y_step when enabled, else the
// platform height alone; a reversed range is swapped.
YLevels y_levels(int32_t enabled, int32_t platform_y, int32_t y_min, int32_t y_max,
                 int32_t y_step);

struct RefineSpec {
    int64_t seed = 0;
    int32_t scan_xz = 1;  // 81 AFK points, else the center only
    YLevels levels;
    int32_t rule = kRuleKeepCenter;
    int32_t pack = 0;     // obs_count = obs | (y + 1024) << 20
};

// Refines rows[0..count) in place (host memory). Returns 0, or -1 on a
// device error (rows unchanged).
int refine_rows(const RefineSpec& spec, ResultRow* rows, int32_t count);

// The same with a world seed per row (spec.seed unused): rows of many worlds
// in one launch.
int refine_rows_multi(const RefineSpec& spec, const int64_t* row_seeds, ResultRow* rows,
                      int32_t count);

// The same as refine_rows without allocating anything: the rows are staged in
// a static device array of kNoAllocMaxRows rows per device (set up at the
// first call on each device). Returns 0, -1 on a device error (rows
// unchanged) and -2 if count exceeds kNoAllocMaxRows.
constexpr int32_t kNoAllocMaxRows = 65536;
int refine_rows_noalloc(const RefineSpec& spec, ResultRow* rows, int32_t count);

// Device time (ns, HIP events) of the last refinement kernel, and the number
// of (row, point, level) evaluations it made.
void refine_last_stats(int64_t* kernel_ns, int64_t* evaluations);

void release_refine();

}  // namespace slime
