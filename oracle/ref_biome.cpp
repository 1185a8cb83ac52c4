// Exposes the reference's terrain spline, compiled from the reference's own
// gpu_noise/spline_generated.cuh (included where it lies) for the CPU, so the
// tests can check this build's data-driven spline evaluation against it. (The
// reference's tree table, gpu_noise/tables/btree262.h, ends in a stray
// character sequence and does not compile as is; the tests compare it as
// text instead.)
// TEST INFRASTRUCTURE ONLY: built by `make -C oracle ref` into oracle/_ref/.
// The CUDA qualifiers of the spline source are defined empty on the compiler
// command line (oracle/Makefile).
#include <cstddef>
#include <cstdint>

#include "spline_generated.cuh"

extern "C" {

float ref_depth_spline(const float* v) { return mcgpu_depth_spline(v); }

}  // extern "C"
