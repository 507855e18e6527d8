// Register-resident butterfly peaks: the VALU ceilings the NTT passes are
// measured against.  Every lane runs independent chains of the butterflies
// the NTT kernels use, with per-lane twiddles held in VGPRs, and touches
// memory only to publish one word, so the rate is bounded by VALU issue
// alone.  One schedule per kind of pass: an iteration applies the stage
// sequence of one 8-bit pass (the 2^16 passes) to each chain, so a kernel's
// peak counts exactly the conditional subtractions, final-stage work and
// normalization its own stages carry.
#include <hip/hip_runtime.h>

#include "../../include/axhel_hip_ubench.h"
#include "arith.hpp"
#include "common.hpp"
#include "ntt_kernels.hpp"

namespace {

using axhel_dev::FwdStage;
using axhel_dev::ModConsts;

constexpr int kChains = 8;  // independent butterflies in flight per lane
constexpr int kStages = 8;  // stages per iteration: one 8-bit pass

template <int S>
__device__ __forceinline__ void pass_stages(uint64_t &x, uint64_t &y, uint64_t w, uint64_t wq, const ModConsts &m,
                                            uint64_t ni, uint64_t niq) {
    if constexpr (S == AXHEL_HIP_UBENCH_FWD_EXACT) {
#pragma unroll
        for (int s = 0; s < kStages; ++s) axhel_dev::fwd_bfly(x, y, w, wq, m);
    } else if constexpr (S == AXHEL_HIP_UBENCH_FWD_RELAXED) {
        // the stage kinds of an 8-bit first (COL) pass
#pragma unroll
        for (int p = kStages - 1; p >= 0; --p) {
            if (axhel_dev::stage_kind<kStages, true>(p) == FwdStage::SUB8)
                axhel_dev::fwd_bfly_relaxed<FwdStage::SUB8>(x, y, w, wq, m);
            else
                axhel_dev::fwd_bfly_relaxed<FwdStage::SKIP>(x, y, w, wq, m);
        }
    } else if constexpr (S == AXHEL_HIP_UBENCH_FWD_RELAXED_LAST) {
        // the stage kinds of an 8-bit last (ROW) pass, then the normalization
#pragma unroll
        for (int p = kStages - 1; p >= 1; --p) {
            if (axhel_dev::stage_kind<kStages, false>(p) == FwdStage::SUB8)
                axhel_dev::fwd_bfly_relaxed<FwdStage::SUB8>(x, y, w, wq, m);
            else
                axhel_dev::fwd_bfly_relaxed<FwdStage::SKIP>(x, y, w, wq, m);
        }
        axhel_dev::fwd_bfly_relaxed<FwdStage::LAST>(x, y, w, wq, m);
        x = axhel_dev::csub_n(axhel_dev::csub_n(x, m.q2, m.nq2), m.q, m.nq);
        y = axhel_dev::csub_n(axhel_dev::csub_n(y, m.q2, m.nq2), m.q, m.nq);
    } else if constexpr (S == AXHEL_HIP_UBENCH_INV) {
#pragma unroll
        for (int s = 0; s < kStages; ++s) axhel_dev::inv_bfly(x, y, w, wq, m);
    } else {  // inverse pass holding the scaled final stage, normalized (CS q) or lazy
#pragma unroll
        for (int s = 0; s < kStages - 1; ++s) axhel_dev::inv_bfly(x, y, w, wq, m);
        axhel_dev::inv_final_bfly(x, y, ni, niq, w, wq, m);
        if constexpr (S == AXHEL_HIP_UBENCH_INV_FINAL_NORM) {
            x = axhel_dev::csub_n(x, m.q, m.nq);
            y = axhel_dev::csub_n(y, m.q, m.nq);
        }
    }
}

template <int S>
__global__ void __launch_bounds__(256) bfly_peak_kernel(uint64_t *out, uint64_t q, uint32_t iters) {
    const ModConsts m = axhel_dev::mod_consts(q);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x[kChains], y[kChains];
    // per-lane twiddles and quotients: any 64-bit words exercise the same
    // instructions (the values are not checked, only timed)
    const uint64_t w = (q >> 3) + gid;
    const uint64_t wq = 0x9E3779B97F4A7C15ull ^ ((uint64_t)gid << 17);
    const uint64_t ni = (q >> 5) + 3 * gid;
    const uint64_t niq = 0x2545F4914F6CDD1Dull ^ ((uint64_t)gid << 9);
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        x[c] = (gid * 2654435761u + c) % q;
        y[c] = (gid * 40503u + 7 * c) % q;
    }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) pass_stages<S>(x[c], y[c], w, wq, m, ni, niq);
    }
    uint64_t acc = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc ^= x[c] + y[c];
    out[gid] = acc;
}

template <int S>
hipError_t launch_peak(uint32_t blocks, uint64_t *out, uint64_t q, uint32_t iters) {
    hipLaunchKernelGGL(bfly_peak_kernel<S>, dim3(blocks), dim3(256), 0, nullptr, out, q, iters);
    return hipGetLastError();
}

hipError_t launch_schedule(int schedule, uint32_t blocks, uint64_t *out, uint64_t q, uint32_t iters) {
    switch (schedule) {
        case AXHEL_HIP_UBENCH_FWD_EXACT: return launch_peak<AXHEL_HIP_UBENCH_FWD_EXACT>(blocks, out, q, iters);
        case AXHEL_HIP_UBENCH_FWD_RELAXED: return launch_peak<AXHEL_HIP_UBENCH_FWD_RELAXED>(blocks, out, q, iters);
        case AXHEL_HIP_UBENCH_FWD_RELAXED_LAST:
            return launch_peak<AXHEL_HIP_UBENCH_FWD_RELAXED_LAST>(blocks, out, q, iters);
        case AXHEL_HIP_UBENCH_INV: return launch_peak<AXHEL_HIP_UBENCH_INV>(blocks, out, q, iters);
        case AXHEL_HIP_UBENCH_INV_FINAL: return launch_peak<AXHEL_HIP_UBENCH_INV_FINAL>(blocks, out, q, iters);
        case AXHEL_HIP_UBENCH_INV_FINAL_NORM:
            return launch_peak<AXHEL_HIP_UBENCH_INV_FINAL_NORM>(blocks, out, q, iters);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

extern "C" axhel_hip_status axhel_hip_ubench_schedule_peak(int schedule, double *butterflies_per_second,
                                                          double *seconds) {
    if (!butterflies_per_second || schedule < 0 || schedule >= AXHEL_HIP_UBENCH_SCHEDULES)
        return AXHEL_HIP_INVALID_ARGUMENT;
    const int cus = axhel_host::cu_count();
    const uint32_t blocks = (uint32_t)cus * 32;  // 8 waves per SIMD, four rounds
    const uint32_t threads = 256;
    const uint32_t iters = 256;                  // x 8 stages
    const uint64_t q = 0x0FFFFFFFFFFC0001ull;    // any 60-bit odd word
    uint64_t *out = nullptr;
    AXHEL_HIP_TRY(hipMalloc(&out, (size_t)blocks * threads * sizeof(uint64_t)));
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = launch_schedule(schedule, blocks, out, q, 8u);  // warm
    if (e == hipSuccess) e = hipEventRecord(a, nullptr);
    if (e == hipSuccess) e = launch_schedule(schedule, blocks, out, q, iters);
    if (e == hipSuccess) e = hipEventRecord(b, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(out);
    AXHEL_HIP_TRY(e);
    const double bflies = (double)blocks * threads * iters * kChains * kStages;
    *butterflies_per_second = bflies / (ms * 1e-3);
    if (seconds) *seconds = ms * 1e-3;
    return AXHEL_HIP_OK;
}

extern "C" axhel_hip_status axhel_hip_ubench_butterfly_peak(double *butterflies_per_second, double *seconds) {
    return axhel_hip_ubench_schedule_peak(AXHEL_HIP_UBENCH_FWD_EXACT, butterflies_per_second, seconds);
}
