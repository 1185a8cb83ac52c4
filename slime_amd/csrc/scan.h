// Device-facing types and launchers of the cluster scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace slime {

// One emitted center: exact 221-cell count, chunk coordinates and the tag of
// the search it belongs to (0 for a single search).
struct Hit {
    int32_t size;
    int32_t cx;
    int32_t cz;
    int32_t tag;
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

// ---- banded first-draw scan -------------------------------------------
//
// Chunks 2^17 apart along x or z share the low 16 bits of the LCG seed sum
// and have opposite first-draw parity. One lane therefore evaluates a slot
// (x, s) for 2 x-phases {x, x+P} and nb z-bands {z0 + b*P + s}: a shared
// low-half computation and one divisibility test per band.
constexpr int32_t kPeriod = 1 << 17;                    // P
constexpr int kBandTileOut = 496;                       // centers per banded x-tile
constexpr int32_t kTilesPerPhase = (kPeriod + kBandTileOut - 1) / kBandTileOut;  // 265
constexpr int kMaxBands = 14;  // three waves per SIMD without spills

struct BandParams {
    int64_t seed;
    int32_t x0;           // first center column of x-block 0, phase 0
    int32_t w;            // centers at offsets >= w from x0 are outside
    int32_t ox_min;       // centers at offsets < ox_min from x0 are outside
    int32_t slot_lo;      // first slot column of the x-tiles (0, or the tail block's)
    int32_t phase_cols;   // slot columns per phase: P with two phases, w with one
    int32_t tiles_per_block;  // x-tiles per block: ceil((phase_cols - slot_lo) / 496)
    int32_t z0;           // center row of band 0, slot 0
    int32_t h_valid;      // center (b, s) is inside iff b*P + s < h_valid
    int32_t s_begin;      // first slot row of this launch (multiple of tile_h)
    int32_t s_end;        // one past the last slot row of this launch
    int32_t tile_h;       // slot rows per tile, multiple of 4
    int32_t tiles_x;      // x-blocks * tiles_per_block
    int32_t tiles_total;
    int32_t min_size, max_size, emit_min;
    int32_t use_annulus;
    int64_t rd_min_sq;
    int32_t acx, acz;
    int32_t tag;          // written into every hit of this job
    Hit* hits;
    uint32_t hit_cap;
    unsigned long long* emitted;
    unsigned long long* hist;
};
// One workgroup per tile; nb bands (1..kMaxBands), ph x-phases (1 or 2; or 3
// with 2..5 bands and the union bound); union_bound selects the rolling
// union-of-disks bound (for nb up to union_bound_max_bands(ph); the square
// bound otherwise).
hipError_t launch_scan_band(const BandParams& p, int nb, int ph, bool union_bound,
                            hipStream_t stream);
// The union-bound kernels exist up to 8 bands, and with two x-phases up to
// 14: counting on the VALU, a thread counts one phase of one band up to 8
// bands, so all four waves share the words (with more, a thread took both
// phases and the full world's 10-band launches ran slower, 1.73 s, than 8-band
// ones, 1.56 s); with the window sums on the matrix cores, a lane holds one
// band-phase row of each of two 16-row operand sets above 8 bands.
// Three x-phases (x, x + P, x + 2P) exist for 2..5 bands with the union bound
// only: 3 NB band-phases, a thread group of 16 each.
constexpr int kUnionBoundMaxBands = 8;
constexpr int kUnionBoundMaxBands2 = 14;
constexpr int kThreePhaseMinBands = 2, kThreePhaseMaxBands = 5;
constexpr int union_bound_max_bands(int ph) {
    return ph == 2 ? kUnionBoundMaxBands2 : ph == 3 ? kThreePhaseMaxBands : kUnionBoundMaxBands;
}
constexpr int min_bands(int ph) { return ph == 3 ? kThreePhaseMinBands : 1; }

// ---- multi-search sweep -----------------------------------------------------
//
// Jobs of many searches (each a BandParams with its own outputs) scanned by
// one persistent launch per kernel class: every workgroup claims tiles from
// per-job cursors in the given job order, skipping jobs whose request is not
// active, until the order is exhausted or the launch's claim budget is spent.
struct SweepClass {
    int nb, ph;
};
constexpr int kSweepClasses = 17;
constexpr SweepClass kSweepClass[kSweepClasses] = {
    {1, 1}, {2, 1}, {3, 1}, {4, 1}, {1, 2}, {2, 2}, {3, 2}, {4, 2}, {6, 2},
    {8, 2}, {12, 2}, {2, 3}, {3, 3}, {4, 3}, {5, 3}, {7, 2}, {14, 2}};
// Class index of (nb, ph), or -1.
inline int sweep_class_of(int nb, int ph) {
    for (int c = 0; c < kSweepClasses; ++c)
        if (kSweepClass[c].nb == nb && kSweepClass[c].ph == ph) return c;
    return -1;
}
// The kernels, each a class with the square bound or the union bound: the
// square bound for the one- and two-phase classes up to 12 bands; the union
// bound for every class but (12, 2) before the wide classes (7, 12 and 14
// bands, two phases: a served full-world request takes them as separate
// calls do) were added; three-phase classes have the union bound only.
struct SweepKernelDef {
    int cls;
    bool ub;
};
constexpr int kSweepKernels = 28;
constexpr SweepKernelDef kSweepKernel[kSweepKernels] = {
    {0, false},  {1, false},  {2, false}, {3, false}, {4, false}, {5, false}, {6, false},
    {7, false},  {8, false},  {9, false}, {10, false}, {0, true}, {1, true},  {2, true},
    {3, true},   {4, true},   {5, true},  {6, true},   {7, true}, {8, true},  {9, true},
    {11, true},  {12, true},  {13, true}, {14, true},  {10, true}, {15, true}, {16, true}};
constexpr int sweep_class_of_kernel(int k) { return kSweepKernel[k].cls; }
constexpr bool sweep_kernel_union(int k) { return kSweepKernel[k].ub; }
constexpr int sweep_kernel_find(int cls, bool ub) {
    for (int k = 0; k < kSweepKernels; ++k)
        if (kSweepKernel[k].cls == cls && kSweepKernel[k].ub == ub) return k;
    return -1;
}
// Whether class c has a kernel with the union bound (ub) or the square bound.
constexpr bool sweep_class_has(int c, bool ub) { return sweep_kernel_find(c, ub) >= 0; }
// The kernel of class c: with the union bound where asked and the class has
// it, else the square bound (-1 for a class without either).
inline int sweep_kernel_of(int cls, bool union_bound) {
    const int k = union_bound ? sweep_kernel_find(cls, true) : -1;
    return k >= 0 ? k : sweep_kernel_find(cls, false);
}

constexpr int32_t kReqActive = 1;  // request status: tiles may be claimed

struct SweepJob {
    BandParams p;
    int32_t cls;  // kernel index (sweep_kernel_of)
    int32_t req;  // request slot whose status gates the job
};
struct SweepArgs {
    const SweepJob* jobs;
    const int32_t* order;   // job indices in claim order
    int32_t norder;
    unsigned int* cursor;   // per job: tiles claimed so far
    unsigned int* done;     // per job: tiles finished so far
    const int32_t* status;  // per request slot
    unsigned int* claims;   // claim attempts in this launch (cleared before it)
    uint32_t budget;        // claim attempts allowed in this launch
};
// A launch without claims: one workgroup per tile, job order[i] running its
// tiles start[i] + (0 .. first[i+1] - first[i] - 1) as the blocks
// [first[i], first[i+1]).
struct BatchArgs {
    const SweepJob* jobs;
    const int32_t* order;
    const int64_t* first;  // n + 1 entries, first[0] == 0
    const int32_t* start;  // n entries
    int32_t n;
    unsigned int* done;    // per job: tiles finished
};
hipError_t launch_batch(const BatchArgs& a, int kernel, int64_t blocks, hipStream_t stream);
// Workgroups of class cls's sweep kernel that fit one CU at once.
int sweep_blocks_per_cu(int kernel);
// Workgroups of the one-tile batch kernel of a kernel index resident per CU.
int batch_blocks_per_cu(int kernel);
// The persistent kernel of class cls over a.order (jobs of that class only).
hipError_t launch_sweep(const SweepArgs& a, int kernel, int grid, hipStream_t stream);
// Sets n words to v with a kernel (usable inside a captured graph).
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t stream);
// Copies the hits of hits[0, n) whose size is at least min_size into out (at
// most cap of them, in no particular order); *count, cleared by the caller,
// ends as the number that qualified (which may exceed cap).
hipError_t launch_select_hits(const Hit* hits, int64_t n, int32_t min_size, Hit* out,
                              uint32_t cap, unsigned int* count, hipStream_t stream);

// ---- first-draw mismatches (Java's nextInt(10) rejection) ------------------
//
// The banded scan counts first draws (next(31) % 10 == 0). It differs from
// Java only where the first draw is rejected (next(31) >= 2^31-8). The seed
// sums s (mod 2^48) of those chunks form a fixed sorted table; per chunk row
// the few candidates are inverted to their x and the affected centers are
// recounted exactly.
struct MismatchChunk {
    int32_t x, z;
};
// Mismatch chunks one search may hold (about 7.5e-10 per chunk: ~10^4 in a
// full world).
constexpr uint32_t kMismatchCap = 1u << 22;
struct CenterFix {
    int64_t seed;
    int32_t cx, cz;
    int32_t first_draw;  // 221-cell count from first draws (what the scan saw)
    int32_t exact;       // Java's count
};
// Sorted table of scrambled-before seed sums s whose first-draw verdict is
// wrong (209,712 entries). Built once on the host.
const std::vector<uint64_t>& mismatch_table();
// Finds every mismatch chunk with z in [z_first, z_first+rows) and x in
// [x_first, x_first+cols) (64-bit extents: cols, rows <= 2^32).
hipError_t launch_find_mismatches(int64_t seed, int32_t x_first, int64_t cols, int32_t z_first,
                                  int64_t rows, const uint64_t* table, int32_t table_n,
                                  MismatchChunk* out, uint32_t cap, unsigned int* count,
                                  hipStream_t stream);
// Mismatch search of many regions at once: query q covers rows
// [row_begin, row_begin + rows) of a virtual row list (ascending row_begin).
struct MismatchQuery {
    int64_t seed;
    int32_t x_first, z_first;
    int64_t cols, rows;
    int64_t row_begin;
    int32_t tag;
};
struct TaggedMismatch {
    int32_t x, z, tag;
};
hipError_t launch_find_mismatches_multi(const MismatchQuery* q, int32_t nq, int64_t total_rows,
                                        const uint64_t* table, int32_t table_n,
                                        TaggedMismatch* out, uint32_t cap, unsigned int* count,
                                        hipStream_t stream);
// Fills first_draw and exact of n centers (each with its own seed).
hipError_t launch_recount(CenterFix* fixes, int32_t n, hipStream_t stream);

}  // namespace slime
