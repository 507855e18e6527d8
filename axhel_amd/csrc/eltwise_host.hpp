// Host-side entry into the batched element-wise kernels (eltwise.hip), used
// by the serving loop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "arith.hpp"

namespace axhel_host {

// Element-wise op codes of the batched kernel.
enum EltBatchOp : int {
    ELT_ADD_VV = 0,
    ELT_ADD_VS = 1,
    ELT_SUB_VV = 2,
    ELT_SUB_VS = 3,
    ELT_MUL = 4,
    ELT_FMA = 5,
    ELT_FMA_NULL = 6,
    ELT_RED_2_1 = 7,
    ELT_RED_4_1 = 8,
    ELT_RED_4_2 = 9,
    ELT_RED_BARRETT = 10,
};

// One limb (vector of n words, one modulus) of a batched element-wise launch.
struct EltLimbDev {
    uint64_t *r;
    const uint64_t *a;
    const uint64_t *b;
    uint64_t s;  // scalar: raw for add/sub vs, ReduceInputNative<F>-reduced for FMA
    axhel_dev::Barrett bc;
};

axhel_dev::Barrett barrett_for(uint64_t q);

// `count` limbs of op/f from the device table d_limbs, each n words.
hipError_t launch_elt_batch(int op, int f, const EltLimbDev *d_limbs, uint32_t count, uint64_t n, hipStream_t st);

}  // namespace axhel_host
