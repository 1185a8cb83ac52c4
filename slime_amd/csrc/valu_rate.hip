// Integer VALU issue rate of this device, the ceiling the scan's roofline is
// read against: a stream of independent 32-bit integer instructions of one
// kind (8 dependency chains per lane, 32 waves per CU), timed with HIP events.
// No counterpart in the reference; measurement only.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/slimefinder.h"

namespace {

constexpr int kRateThreads = 256;
constexpr int kRateBlocksPerCu = 8;  // 32 waves per CU: 8 per SIMD
constexpr int kOpsPerIter = 64;      // 8 chains x 8 instructions

#define CHAIN8(INSN)                                                                     \
    asm volatile(INSN " %0, %0, %8\n\t" INSN " %1, %1, %8\n\t" INSN " %2, %2, %8\n\t" INSN \
                      " %3, %3, %8\n\t" INSN " %4, %4, %8\n\t" INSN " %5, %5, %8\n\t" INSN \
                      " %6, %6, %8\n\t" INSN " %7, %7, %8"                               \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6),    \
                   "+v"(a7)                                                               \
                 : "v"(k))

#define CHAIN8_3(INSN)                                                                      \
    asm volatile(INSN " %0, %0, %8, %0\n\t" INSN " %1, %1, %8, %1\n\t" INSN                 \
                      " %2, %2, %8, %2\n\t" INSN " %3, %3, %8, %3\n\t" INSN                 \
                      " %4, %4, %8, %4\n\t" INSN " %5, %5, %8, %5\n\t" INSN                 \
                      " %6, %6, %8, %6\n\t" INSN " %7, %7, %8, %7"                          \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6),    \
                   "+v"(a7)                                                                 \
                 : "v"(k))

#define CHAIN8_LIT(INSN)                                                                  \
    asm volatile(INSN " %0, 0x9e3779b9, %0\n\t" INSN " %1, 0x9e3779b9, %1\n\t" INSN         \
                      " %2, 0x9e3779b9, %2\n\t" INSN " %3, 0x9e3779b9, %3\n\t" INSN         \
                      " %4, 0x9e3779b9, %4\n\t" INSN " %5, 0x9e3779b9, %5\n\t" INSN         \
                      " %6, 0x9e3779b9, %6\n\t" INSN " %7, 0x9e3779b9, %7"                    \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6),    \
                   "+v"(a7))

#define CHAIN8_SGPR(INSN)                                                                   \
    asm volatile(INSN " %0, %8, %0\n\t" INSN " %1, %8, %1\n\t" INSN " %2, %8, %2\n\t" INSN \
                      " %3, %8, %3\n\t" INSN " %4, %8, %4\n\t" INSN " %5, %8, %5\n\t" INSN \
                      " %6, %8, %6\n\t" INSN " %7, %8, %7"                                  \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6),    \
                   "+v"(a7)                                                                 \
                 : "s"(ks))

// kind: 0 v_add_u32, 1 v_xor_b32 (VOP2, 4 bytes); 2 v_mul_lo_u32,
// 3 v_bcnt_u32_b32, 4 v_mad_u32_u24, 5 v_alignbit_b32 (VOP3, 8 bytes);
// 6 v_xor_b32 with a 32-bit literal (VOP2 + literal, 8 bytes); 7 v_add_u32
// in its VOP3 encoding (8 bytes); 8 v_mul_u32_u24 (VOP2, 4 bytes);
// 9 v_xor_b32 with an SGPR operand (VOP2, 4 bytes).
template <int KIND>
__global__ __launch_bounds__(kRateThreads) void valu_rate_kernel(uint32_t* out, int iters,
                                                                 uint32_t seed) {
    uint32_t a0 = threadIdx.x ^ seed, a1 = a0 * 3u, a2 = a0 * 5u, a3 = a0 * 7u, a4 = a0 * 11u,
             a5 = a0 * 13u, a6 = a0 * 17u, a7 = a0 * 19u;
    const uint32_t k = seed | 1u;
    const uint32_t ks = __builtin_amdgcn_readfirstlane(k);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kOpsPerIter / 8; ++u) {
            if constexpr (KIND == 0) CHAIN8("v_add_u32");
            else if constexpr (KIND == 1) CHAIN8("v_xor_b32");
            else if constexpr (KIND == 2) CHAIN8("v_mul_lo_u32");
            else if constexpr (KIND == 3) CHAIN8("v_bcnt_u32_b32");
            else if constexpr (KIND == 4) CHAIN8_3("v_mad_u32_u24");
            else if constexpr (KIND == 5) CHAIN8_3("v_alignbit_b32");
            else if constexpr (KIND == 6) CHAIN8_LIT("v_xor_b32");
            else if constexpr (KIND == 7) CHAIN8("v_add_u32_e64");
            else if constexpr (KIND == 8) CHAIN8("v_mul_u32_u24");
            else CHAIN8_SGPR("v_xor_b32");
        }
    }
    const uint32_t r = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
    if (r == 0x9E3779B9u) out[blockIdx.x] = r;  // keeps the chains live; practically never taken
}

#undef CHAIN8
#undef CHAIN8_3
#undef CHAIN8_LIT
#undef CHAIN8_SGPR

hipError_t launch_rate(int kind, int blocks, uint32_t* out, int iters, hipStream_t s) {
    switch (kind) {
        case 0: hipLaunchKernelGGL(valu_rate_kernel<0>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 1: hipLaunchKernelGGL(valu_rate_kernel<1>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 2: hipLaunchKernelGGL(valu_rate_kernel<2>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 3: hipLaunchKernelGGL(valu_rate_kernel<3>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 4: hipLaunchKernelGGL(valu_rate_kernel<4>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 5: hipLaunchKernelGGL(valu_rate_kernel<5>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 6: hipLaunchKernelGGL(valu_rate_kernel<6>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 7: hipLaunchKernelGGL(valu_rate_kernel<7>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        case 8: hipLaunchKernelGGL(valu_rate_kernel<8>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
        default: hipLaunchKernelGGL(valu_rate_kernel<9>, dim3(blocks), dim3(kRateThreads), 0, s, out, iters, 0x2545F491u); break;
    }
    return hipGetLastError();
}

}  // namespace

extern "C" int32_t slime_valu_rate(int32_t kind, int32_t iters, int64_t* lane_instructions,
                                   int64_t* kernel_ns) {
    if (kind < 0 || kind > 9 || iters <= 0 || iters > (1 << 20)) return -1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
        (void)hipGetLastError();
        return -1;
    }
    const int blocks = cus * kRateBlocksPerCu;
    uint32_t* out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    bool ok = hipMalloc(&out, sizeof(uint32_t) * (size_t)blocks) == hipSuccess &&
              hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              launch_rate(kind, blocks, out, 16, s) == hipSuccess &&  // warm-up
              hipEventRecord(e0, s) == hipSuccess &&
              launch_rate(kind, blocks, out, iters, s) == hipSuccess &&
              hipEventRecord(e1, s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess &&
              hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    if (!ok) {
        (void)hipGetLastError();
        return -1;
    }
    if (lane_instructions)
        *lane_instructions = (int64_t)blocks * kRateThreads * (int64_t)iters * kOpsPerIter;
    if (kernel_ns) *kernel_ns = (int64_t)((double)ms * 1e6);
    return 0;
}
