// Banded scan kernels with one x-phase, for regions at most 2^17 columns wide
// (the tile is band_tile.h).
#include "band_launch.h"

namespace slime {
namespace band {
template hipError_t launch_phases<1>(const BandParams&, int, bool, hipStream_t);
}  // namespace band
}  // namespace slime
