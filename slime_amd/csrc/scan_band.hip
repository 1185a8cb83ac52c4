// Banded scan kernels with two x-phases (regions wider than 2^17 columns) and
// the launch entry of both forms; the tile itself is band_tile.h.
#include "band_launch.h"

namespace slime {

namespace band {
template hipError_t launch_phases<2>(const BandParams&, int, bool, hipStream_t);
}  // namespace band

hipError_t launch_scan_band(const BandParams& p, int nb, int ph, bool union_bound,
                            hipStream_t stream) {
    if (p.tiles_total <= 0) return hipSuccess;
    if (ph == 2) return band::launch_phases<2>(p, nb, union_bound, stream);
    if (ph == 1) return band::launch_phases<1>(p, nb, union_bound, stream);
    if (ph == 3) return band::launch_phases3(p, nb, union_bound, stream);
    return hipErrorInvalidValue;
}

}  // namespace slime
