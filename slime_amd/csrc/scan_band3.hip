// Banded scan kernels with three x-phases (x, x + 2^17, x + 2^18: regions
// between 2^18 and 3 * 2^17 columns wide are covered without masked columns);
// the tile is band_tile.h.
#include "band_launch.h"

namespace slime {
namespace band {

hipError_t launch_phases3(const BandParams& p, int nb, bool ub, hipStream_t stream) {
    if (!ub) return hipErrorInvalidValue;  // the union bound only
    switch (nb) {
        case 2: return launch_three<2>(p, stream);
        case 3: return launch_three<3>(p, stream);
        case 4: return launch_three<4>(p, stream);
        case 5: return launch_three<5>(p, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace band
}  // namespace slime
