// Per-search banded kernels: one workgroup per tile of one job.
#pragma once

#include "band_tile.h"

namespace slime {
namespace band {

// Three waves per SIMD (<= 168 VGPRs): enough to overlap the in-order VALU
// stream of the generation with the LDS waits of the counting.
template <int NB, int PH, bool UB>
__global__ __launch_bounds__(kThreads, 3) void scan_band_kernel(BandParams p) {
    __shared__ TileLds<NB, PH> lds;
    for (int i = threadIdx.x; i < kHistBins; i += kThreads) lds.lhist[i] = 0;
    band_tile<NB, PH, UB>(p, p.emit_min, blockIdx.x, lds, (int)threadIdx.x);
    flush_hist(lds.lhist, p.hist);
}

template <int NB, int PH>
hipError_t launch_one(const BandParams& p, bool ub, hipStream_t stream) {
    constexpr bool kUnion = NB <= union_bound_max_bands(PH);
    if (ub && kUnion)
        hipLaunchKernelGGL((scan_band_kernel<NB, PH, kUnion>), dim3(p.tiles_total),
                           dim3(kThreads), 0, stream, p);
    else
        hipLaunchKernelGGL((scan_band_kernel<NB, PH, false>), dim3(p.tiles_total),
                           dim3(kThreads), 0, stream, p);
    return hipGetLastError();
}

template <int PH>
hipError_t launch_phases(const BandParams& p, int nb, bool ub, hipStream_t stream) {
    switch (nb) {
        case 1: return launch_one<1, PH>(p, ub, stream);
        case 2: return launch_one<2, PH>(p, ub, stream);
        case 3: return launch_one<3, PH>(p, ub, stream);
        case 4: return launch_one<4, PH>(p, ub, stream);
        case 5: return launch_one<5, PH>(p, ub, stream);
        case 6: return launch_one<6, PH>(p, ub, stream);
        case 7: return launch_one<7, PH>(p, ub, stream);
        case 8: return launch_one<8, PH>(p, ub, stream);
        case 9: return launch_one<9, PH>(p, ub, stream);
        case 10: return launch_one<10, PH>(p, ub, stream);
        case 11: return launch_one<11, PH>(p, ub, stream);
        case 12: return launch_one<12, PH>(p, ub, stream);
        case 13: return launch_one<13, PH>(p, ub, stream);
        case 14: return launch_one<14, PH>(p, ub, stream);
        default: return hipErrorInvalidValue;
    }
}

// Three x-phases: 2..5 bands, union bound only (scan_band3.hip).
template <int NB>
hipError_t launch_three(const BandParams& p, hipStream_t stream) {
    hipLaunchKernelGGL((scan_band_kernel<NB, 3, true>), dim3(p.tiles_total), dim3(kThreads), 0,
                       stream, p);
    return hipGetLastError();
}
hipError_t launch_phases3(const BandParams& p, int nb, bool ub, hipStream_t stream);

// Instantiated in scan_band.hip (two phases) and scan_band1.hip (one phase).
extern template hipError_t launch_phases<1>(const BandParams&, int, bool, hipStream_t);
extern template hipError_t launch_phases<2>(const BandParams&, int, bool, hipStream_t);

}  // namespace band
}  // namespace slime
