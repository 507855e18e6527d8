// Element-wise modular kernels for gfx950 and their C-ABI entry points.
//
// One streaming kernel template covers the five reference functions
// (axhel/eltwise/eltwise-{add,sub,mul,fma,reduce}-mod.cpp).  The work is
// HBM-bound: each thread moves UNROLL 16-byte vectors per operand with all
// loads issued before any arithmetic, so enough bytes are in flight to cover
// HBM latency.  Pointers that are only 8-byte aligned (legal in the reference,
// tests/public/test_eltwise.cpp:518-561) take the 8-byte variant of the same
// kernel.  The per-element arithmetic is the reference's, word for word.
#include <hip/hip_runtime.h>

#include "arith.hpp"
#include "common.hpp"
#include "eltwise_host.hpp"

using axhel_dev::Barrett;

namespace {

enum EltOp : int {
    OP_ADD_VV,
    OP_ADD_VS,
    OP_SUB_VV,
    OP_SUB_VS,
    OP_MUL,
    OP_FMA,
    OP_FMA_NULL,
    OP_RED_2_1,
    OP_RED_4_1,
    OP_RED_4_2,
    OP_RED_BARRETT,
};

constexpr int kBlock = 256;
constexpr int kUnroll = 4;

// Operands stream through exactly once: non-temporal loads and stores keep
// them from churning L2 (about 12 % more HBM bandwidth on gfx950 than plain
// accesses for these kernels).
// Global-space (address_space 1) pointers: the accesses issue as global_*
// instructions rather than flat_*.
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) uint64_t gword;
typedef __attribute__((address_space(1))) u64x2 gword2;
__device__ __forceinline__ ulonglong2 ldv(const ulonglong2 *p) {
    const u64x2 x = __builtin_nontemporal_load((const gword2 *)p);
    ulonglong2 r;
    r.x = x.x;
    r.y = x.y;
    return r;
}
__device__ __forceinline__ void stv(ulonglong2 *p, ulonglong2 v) {
    u64x2 x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, (gword2 *)p);
}
__device__ __forceinline__ uint64_t lds1(const uint64_t *p) { return __builtin_nontemporal_load((const gword *)p); }
__device__ __forceinline__ void sts1(uint64_t *p, uint64_t v) { __builtin_nontemporal_store(v, (gword *)p); }

// Whether the op reads a second vector operand.
constexpr bool has_b(int op) { return op == OP_ADD_VV || op == OP_SUB_VV || op == OP_MUL || op == OP_FMA; }

// One element.  `s` is the scalar operand (add/sub vs: raw as in the
// reference; FMA: already reduced by ReduceInputNative<F> on the host,
// eltwise-fma-mod.cpp:32).
template <int OP, int F>
__device__ __forceinline__ uint64_t elt_apply(uint64_t x, uint64_t y, uint64_t s, const Barrett &bc) {
    const uint64_t q = bc.q;
    if constexpr (OP == OP_ADD_VV || OP == OP_ADD_VS) {
        // eltwise-add-mod.cpp:20-26
        const uint64_t v = OP == OP_ADD_VV ? y : s;
        uint64_t sum = x + v;
        sum -= (uint64_t)(sum >= q) * q;
        return sum;
    } else if constexpr (OP == OP_SUB_VV || OP == OP_SUB_VS) {
        // eltwise-sub-mod.cpp:20-26
        const uint64_t v = OP == OP_SUB_VV ? y : s;
        uint64_t d = x - v;
        d += (uint64_t)(x < v) * q;
        return d;
    } else if constexpr (OP == OP_MUL) {
        // eltwise-mul-mod.cpp:34-61
        const uint64_t a = axhel_dev::reduce_input<F>(x, q);
        const uint64_t b = axhel_dev::reduce_input<F>(y, q);
        return axhel_dev::barrett_mul(a, b, bc);
    } else if constexpr (OP == OP_FMA || OP == OP_FMA_NULL) {
        // eltwise-fma-mod.cpp:37-77
        const uint64_t a = axhel_dev::reduce_input<F>(x, q);
        uint64_t v = axhel_dev::barrett_mul(a, s, bc);
        if constexpr (OP == OP_FMA) {
            v += axhel_dev::reduce_input<F>(y, q);
            if (v >= q) v -= q;
        }
        return v;
    } else if constexpr (OP == OP_RED_2_1) {
        return axhel_dev::csub(x, q);  // eltwise-reduce-mod.cpp:55-64
    } else if constexpr (OP == OP_RED_4_1) {
        return axhel_dev::csub(axhel_dev::csub(x, q << 1), q);  // :67-86
    } else if constexpr (OP == OP_RED_4_2) {
        return axhel_dev::csub(x, q << 1);  // :89-105
    } else {
        return axhel_dev::barrett_reduce64(x, bc);  // :108-127
    }
}

// 16-byte path: n2 two-word vectors; a lone trailing word (odd n) is done by
// the first thread of the grid.
template <int OP, int F>
__global__ void __launch_bounds__(kBlock)
elt_kernel_v2(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t s, uint64_t n, Barrett bc) {
    const uint64_t n2 = n >> 1;
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kUnroll) + threadIdx.x;
    const ulonglong2 *a2 = reinterpret_cast<const ulonglong2 *>(a);
    const ulonglong2 *b2 = reinterpret_cast<const ulonglong2 *>(b);
    ulonglong2 *r2 = reinterpret_cast<ulonglong2 *>(r);
    ulonglong2 xa[kUnroll], xb[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        if (i < n2) {
            xa[u] = ldv(a2 + i);
            if constexpr (has_b(OP)) xb[u] = ldv(b2 + i);
        }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        if (i < n2) {
            ulonglong2 o;
            o.x = elt_apply<OP, F>(xa[u].x, has_b(OP) ? xb[u].x : 0, s, bc);
            o.y = elt_apply<OP, F>(xa[u].y, has_b(OP) ? xb[u].y : 0, s, bc);
            stv(r2 + i, o);
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t i = n - 1;
        const uint64_t y = has_b(OP) ? b[i] : 0;
        r[i] = elt_apply<OP, F>(a[i], y, s, bc);
    }
}

// 8-byte path for operands that are not all 16-byte aligned.
template <int OP, int F>
__global__ void __launch_bounds__(kBlock)
elt_kernel_v1(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t s, uint64_t n, Barrett bc) {
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kUnroll * 2) + threadIdx.x;
    constexpr int U = kUnroll * 2;
    uint64_t xa[U], xb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        if (i < n) {
            xa[u] = lds1(a + i);
            if constexpr (has_b(OP)) xb[u] = lds1(b + i);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        if (i < n) sts1(r + i, elt_apply<OP, F>(xa[u], has_b(OP) ? xb[u] : 0, s, bc));
    }
}

// Barrett constants of eltwise-mul-mod.cpp:22-30 / multiply-factor.cpp:11-19.
Barrett make_barrett(uint64_t q) {
    Barrett b{};
    b.q = q;
    if (q < 2) return b;  // callers reject q < 2 on Barrett paths
    const unsigned nbits = 64u - (unsigned)__builtin_clzll(q);
    b.s = nbits - 2;
    b.s_comp = (64u - b.s) & 63u;
    const unsigned __int128 num = ((unsigned __int128)1 << (nbits - 2)) << 64;
    b.mu = (uint64_t)(num / q);
    return b;
}

template <int OP, int F>
axhel_hip_status launch_elt(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t s, uint64_t n,
                            uint64_t q, void *stream) {
    if (n == 0) return AXHEL_HIP_OK;
    if (r == nullptr || a == nullptr || (has_b(OP) && b == nullptr)) return AXHEL_HIP_INVALID_ARGUMENT;
    const hipStream_t st = axhel_host::as_stream(stream);
    axhel_host::Staging stg(st);
    const uint64_t *da = nullptr, *db = nullptr;
    uint64_t *dr = nullptr;
    AXHEL_HIP_TRY(stg.input(a, n, &da));
    if (has_b(OP)) AXHEL_HIP_TRY(stg.input(b, n, &db));
    AXHEL_HIP_TRY(stg.output(r, n, &dr));
    const Barrett bc = make_barrett(q);
    const bool vec = axhel_host::aligned16(da) && axhel_host::aligned16(dr) && (!has_b(OP) || axhel_host::aligned16(db));
    if (vec) {
        const uint64_t n2 = n >> 1;
        uint64_t blocks = (n2 + kBlock * kUnroll - 1) / (kBlock * kUnroll);
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL((elt_kernel_v2<OP, F>), dim3((unsigned)blocks), dim3(kBlock), 0, st, dr, da, db, s, n, bc);
    } else {
        const uint64_t per = (uint64_t)kBlock * kUnroll * 2;
        const uint64_t blocks = (n + per - 1) / per;
        hipLaunchKernelGGL((elt_kernel_v1<OP, F>), dim3((unsigned)blocks), dim3(kBlock), 0, st, dr, da, db, s, n, bc);
    }
    AXHEL_HIP_TRY(hipGetLastError());
    AXHEL_HIP_TRY(stg.finish());
    return AXHEL_HIP_OK;
}

// Device or staged copy for the identity branches of EltwiseReduceMod.
axhel_hip_status copy_words(uint64_t *r, const uint64_t *a, uint64_t n, void *stream) {
    if (n == 0 || r == a) return AXHEL_HIP_OK;  // in-place identity: nothing to move
    const hipStream_t st = axhel_host::as_stream(stream);
    axhel_host::Staging stg(st);
    const uint64_t *da = nullptr;
    uint64_t *dr = nullptr;
    AXHEL_HIP_TRY(stg.input(a, n, &da));
    AXHEL_HIP_TRY(stg.output(r, n, &dr));
    AXHEL_HIP_TRY(hipMemcpyAsync(dr, da, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
    AXHEL_HIP_TRY(stg.finish());
    return AXHEL_HIP_OK;
}

}  // namespace

// ------------------------------------------------------ batched element-wise

// Several limbs of one element-wise op in one launch: limb j covers n words
// with its own modulus; grid = limbs x chunks.  Used by the serving loop for
// RNS requests (one request = `limbs` vectors, one modulus each).
namespace {

static_assert((int)OP_RED_BARRETT == (int)axhel_host::ELT_RED_BARRETT, "op codes shared with eltwise_host.hpp");

constexpr int kBatchChunk = kBlock * kUnroll;  // words per workgroup

template <int OP, int F>
__global__ void __launch_bounds__(kBlock)
elt_batch_kernel(const axhel_host::EltLimbDev *limbs, uint64_t n, uint32_t chunks) {
    const uint32_t j = blockIdx.x / chunks;
    const uint64_t c0 = (uint64_t)(blockIdx.x % chunks) * kBatchChunk;
    const axhel_host::EltLimbDev L = limbs[j];
    uint64_t xa[kUnroll], xb[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t i = c0 + (uint64_t)u * kBlock + threadIdx.x;
        if (i < n) {
            xa[u] = lds1(L.a + i);
            if constexpr (has_b(OP)) xb[u] = lds1(L.b + i);
        }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint64_t i = c0 + (uint64_t)u * kBlock + threadIdx.x;
        if (i < n) sts1(L.r + i, elt_apply<OP, F>(xa[u], has_b(OP) ? xb[u] : 0, L.s, L.bc));
    }
}

template <int OP, int F>
hipError_t launch_batch(const axhel_host::EltLimbDev *d_limbs, uint32_t count, uint64_t n, hipStream_t st) {
    const uint64_t chunks = (n + kBatchChunk - 1) / kBatchChunk;
    const uint64_t blocks = chunks * count;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((elt_batch_kernel<OP, F>), dim3((unsigned)blocks), dim3(kBlock), 0, st, d_limbs, n,
                       (uint32_t)chunks);
    return hipGetLastError();
}

}  // namespace

namespace axhel_host {

Barrett barrett_for(uint64_t q) { return make_barrett(q); }

// op: 0 add_vv, 1 add_vs, 2 sub_vv, 3 sub_vs, 4 mul, 5 fma, 6 fma (null addend),
// 7.. reduce modes (7: 2->1, 8: 4->1, 9: 4->2, 10: Barrett).  f = input factor.
hipError_t launch_elt_batch(int op, int f, const EltLimbDev *d_limbs, uint32_t count, uint64_t n, hipStream_t st) {
#define AXHEL_EB(OPV, FV) return launch_batch<OPV, FV>(d_limbs, count, n, st)
    switch (op) {
        case OP_ADD_VV: AXHEL_EB(OP_ADD_VV, 1);
        case OP_ADD_VS: AXHEL_EB(OP_ADD_VS, 1);
        case OP_SUB_VV: AXHEL_EB(OP_SUB_VV, 1);
        case OP_SUB_VS: AXHEL_EB(OP_SUB_VS, 1);
        case OP_MUL:
            if (f == 1) AXHEL_EB(OP_MUL, 1);
            if (f == 2) AXHEL_EB(OP_MUL, 2);
            AXHEL_EB(OP_MUL, 4);
        case OP_FMA:
            if (f == 1) AXHEL_EB(OP_FMA, 1);
            if (f == 2) AXHEL_EB(OP_FMA, 2);
            if (f == 4) AXHEL_EB(OP_FMA, 4);
            AXHEL_EB(OP_FMA, 8);
        case OP_FMA_NULL:
            if (f == 1) AXHEL_EB(OP_FMA_NULL, 1);
            if (f == 2) AXHEL_EB(OP_FMA_NULL, 2);
            if (f == 4) AXHEL_EB(OP_FMA_NULL, 4);
            AXHEL_EB(OP_FMA_NULL, 8);
        case OP_RED_2_1: AXHEL_EB(OP_RED_2_1, 1);
        case OP_RED_4_1: AXHEL_EB(OP_RED_4_1, 1);
        case OP_RED_4_2: AXHEL_EB(OP_RED_4_2, 1);
        case OP_RED_BARRETT: AXHEL_EB(OP_RED_BARRETT, 1);
    }
#undef AXHEL_EB
    return hipErrorInvalidValue;
}

}  // namespace axhel_host

extern "C" {

axhel_hip_status axhel_hip_add_mod_vv(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                                      uint64_t n, uint64_t modulus, void *stream) {
    return launch_elt<OP_ADD_VV, 1>(result, operand1, operand2, 0, n, modulus, stream);
}

axhel_hip_status axhel_hip_add_mod_vs(uint64_t *result, const uint64_t *operand1, uint64_t operand2, uint64_t n,
                                      uint64_t modulus, void *stream) {
    return launch_elt<OP_ADD_VS, 1>(result, operand1, nullptr, operand2, n, modulus, stream);
}

axhel_hip_status axhel_hip_sub_mod_vv(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                                      uint64_t n, uint64_t modulus, void *stream) {
    return launch_elt<OP_SUB_VV, 1>(result, operand1, operand2, 0, n, modulus, stream);
}

axhel_hip_status axhel_hip_sub_mod_vs(uint64_t *result, const uint64_t *operand1, uint64_t operand2, uint64_t n,
                                      uint64_t modulus, void *stream) {
    return launch_elt<OP_SUB_VS, 1>(result, operand1, nullptr, operand2, n, modulus, stream);
}

axhel_hip_status axhel_hip_mul_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2, uint64_t n,
                                   uint64_t modulus, uint64_t input_mod_factor, void *stream) {
    if (input_mod_factor != 1 && input_mod_factor != 2 && input_mod_factor != 4) return AXHEL_HIP_UNSUPPORTED_FACTOR;
    if (modulus < 2) return AXHEL_HIP_INVALID_ARGUMENT;
    switch (input_mod_factor) {
        case 1: return launch_elt<OP_MUL, 1>(result, operand1, operand2, 0, n, modulus, stream);
        case 2: return launch_elt<OP_MUL, 2>(result, operand1, operand2, 0, n, modulus, stream);
        default: return launch_elt<OP_MUL, 4>(result, operand1, operand2, 0, n, modulus, stream);
    }
}

axhel_hip_status axhel_hip_fma_mod(uint64_t *result, const uint64_t *operand1, uint64_t operand2,
                                   const uint64_t *operand3, uint64_t n, uint64_t modulus, uint64_t input_mod_factor,
                                   void *stream) {
    if (input_mod_factor != 1 && input_mod_factor != 2 && input_mod_factor != 4 && input_mod_factor != 8)
        return AXHEL_HIP_UNSUPPORTED_FACTOR;
    if (modulus < 2) return AXHEL_HIP_INVALID_ARGUMENT;
    // ReduceInputNative<F>(operand2) once, eltwise-fma-mod.cpp:32.
    uint64_t s = operand2;
    if (input_mod_factor == 8 && s >= (modulus << 2)) s -= modulus << 2;
    if (input_mod_factor >= 4 && s >= (modulus << 1)) s -= modulus << 1;
    if (input_mod_factor >= 2 && s >= modulus) s -= modulus;
    const bool add = operand3 != nullptr;
#define AXHEL_FMA_CASE(F)                                                                                 \
    case F:                                                                                               \
        return add ? launch_elt<OP_FMA, F>(result, operand1, operand3, s, n, modulus, stream)            \
                   : launch_elt<OP_FMA_NULL, F>(result, operand1, nullptr, s, n, modulus, stream);
    switch (input_mod_factor) {
        AXHEL_FMA_CASE(1)
        AXHEL_FMA_CASE(2)
        AXHEL_FMA_CASE(4)
        AXHEL_FMA_CASE(8)
    }
#undef AXHEL_FMA_CASE
    return AXHEL_HIP_UNSUPPORTED_FACTOR;
}

axhel_hip_status axhel_hip_reduce_mod(uint64_t *result, const uint64_t *operand, uint64_t n, uint64_t modulus,
                                      uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream) {
    // Branch order of eltwise-reduce-mod.cpp:32-128.
    if (input_mod_factor == 1 || (output_mod_factor == 2 && input_mod_factor <= 2))
        return copy_words(result, operand, n, stream);
    if (input_mod_factor == 2 && output_mod_factor == 1)
        return launch_elt<OP_RED_2_1, 1>(result, operand, nullptr, 0, n, modulus, stream);
    if (input_mod_factor == 4 && output_mod_factor == 1)
        return launch_elt<OP_RED_4_1, 1>(result, operand, nullptr, 0, n, modulus, stream);
    if (input_mod_factor == 4 && output_mod_factor == 2)
        return launch_elt<OP_RED_4_2, 1>(result, operand, nullptr, 0, n, modulus, stream);
    if (input_mod_factor == modulus) {
        if (modulus < 2) return AXHEL_HIP_INVALID_ARGUMENT;
        return launch_elt<OP_RED_BARRETT, 1>(result, operand, nullptr, 0, n, modulus, stream);
    }
    return AXHEL_HIP_UNSUPPORTED_FACTOR;
}

}  // extern "C"
