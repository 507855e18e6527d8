// Register-resident Harvey butterfly peak: the VALU ceiling the NTT passes are
// measured against.  Every lane runs independent chains of the exact forward
// butterfly the NTT kernels use (64-bit Shoup product, conditional subtract,
// two adds) with per-lane twiddles held in VGPRs, and touches memory only to
// publish one word, so the rate is bounded by VALU issue alone.
#include <hip/hip_runtime.h>

#include "../../include/axhel_hip_ubench.h"
#include "arith.hpp"
#include "common.hpp"

namespace {

constexpr int kChains = 8;  // independent butterflies in flight per lane

__global__ void __launch_bounds__(256) bfly_peak_kernel(uint64_t *out, uint64_t q, uint32_t iters) {
    const axhel_dev::ModConsts m = axhel_dev::mod_consts(q);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x[kChains], y[kChains];
    // per-lane twiddle and Shoup quotient (valid pair: w < q, wq = floor(w 2^64 / q)
    // is not needed for timing; any 64-bit words exercise the same instructions)
    const uint64_t w = (q >> 3) + gid;
    const uint64_t wq = 0x9E3779B97F4A7C15ull ^ ((uint64_t)gid << 17);
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        x[c] = (gid * 2654435761u + c) % q;
        y[c] = (gid * 40503u + 7 * c) % q;
    }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) axhel_dev::fwd_bfly(x[c], y[c], w, wq, m);
    }
    uint64_t acc = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc ^= x[c] + y[c];
    out[gid] = acc;
}

}  // namespace

extern "C" axhel_hip_status axhel_hip_ubench_butterfly_peak(double *butterflies_per_second, double *seconds) {
    if (!butterflies_per_second) return AXHEL_HIP_INVALID_ARGUMENT;
    const int cus = axhel_host::cu_count();
    const uint32_t blocks = (uint32_t)cus * 32;  // 8 waves per SIMD
    const uint32_t threads = 256;
    const uint32_t iters = 2048;
    const uint64_t q = 0x0FFFFFFFFFFC0001ull;  // any 60-bit odd word
    uint64_t *out = nullptr;
    AXHEL_HIP_TRY(hipMalloc(&out, (size_t)blocks * threads * sizeof(uint64_t)));
    hipEvent_t a, b;
    AXHEL_HIP_TRY(hipEventCreate(&a));
    AXHEL_HIP_TRY(hipEventCreate(&b));
    hipLaunchKernelGGL(bfly_peak_kernel, dim3(blocks), dim3(threads), 0, nullptr, out, q, 64u);  // warm
    AXHEL_HIP_TRY(hipEventRecord(a, nullptr));
    hipLaunchKernelGGL(bfly_peak_kernel, dim3(blocks), dim3(threads), 0, nullptr, out, q, iters);
    AXHEL_HIP_TRY(hipEventRecord(b, nullptr));
    AXHEL_HIP_TRY(hipEventSynchronize(b));
    float ms = 0;
    AXHEL_HIP_TRY(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(out);
    const double bflies = (double)blocks * threads * iters * kChains;
    *butterflies_per_second = bflies / (ms * 1e-3);
    if (seconds) *seconds = ms * 1e-3;
    return AXHEL_HIP_OK;
}
