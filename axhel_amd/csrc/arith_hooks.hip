// Test hooks of the device arithmetic primitives (arith.hpp): each
// primitive applied element-wise on the GPU, so tests can compare the
// inline-assembly forms with exact integer arithmetic word by word.  The
// reference keeps the same kind of hooks for its vector kernels
// (axhel/ntt/ntt-sve.cpp:1462-1498, exercised by
// tests/internal/test_internal_sve.cpp:254-324) and tests its scalar
// primitives directly (tests/internal/test_internal_native.cpp:26-158).
// Built as its own library (libaxhel_arith_hooks.so); nothing in the
// product calls it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "arith.hpp"

namespace {

using namespace axhel_dev;

// op codes of axhel_arith_hooks_apply; `in` holds 7 words per item
// (a..g below), `out` 2 words per item
enum Op {
    OP_MULHI64 = 0,         // out0 = hi64(a * b)
    OP_SHOUP_ACC = 1,       // out0 = a + shoup(b, c, d, q = e)          (nq = 2^64 - e)
    OP_CSUB_N = 2,          // out0 = csub(a, b)                          (nb = 2^64 - b)
    OP_SUB64 = 3,           // out0 = a - b mod 2^64
    OP_TWICE_PLUS = 4,      // out0 = 2a + b mod 2^64
    OP_FWD_BFLY = 5,        // (out0, out1) = fwd_bfly(x = a, y = b, w = c, wq = d, q = e)
    OP_FWD_SKIP = 6,        // fwd_bfly_relaxed<SKIP>, same operands
    OP_FWD_SUB8 = 7,        // fwd_bfly_relaxed<SUB8>
    OP_FWD_LAST = 8,        // fwd_bfly_relaxed<LAST>
    OP_INV_BFLY = 9,        // inv_bfly(x = a, y = b, w = c, wq = d, q = e)
    OP_INV_FINAL = 10,      // inv_final_bfly(x = a, y = b, ni = c, niq = d, s = f, sq = g, q = e)
    OP_BARRETT_MUL = 11,    // out0 = barrett_mul(a, b) for modulus e
    OP_BARRETT_REDUCE = 12, // out0 = barrett_reduce64(a) for modulus e
    OP_COUNT = 13
};

__global__ void hooks_kernel(int op, const uint64_t *in, uint64_t *out, uint64_t count, uint64_t mu, uint32_t s,
                             uint32_t s_comp) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t *p = in + 7 * i;
    uint64_t a = p[0], b = p[1];
    const uint64_t c = p[2], d = p[3], e = p[4], f = p[5], g = p[6];
    uint64_t r0 = 0, r1 = 0;
    const ModConsts m = mod_consts(e);
    switch (op) {
        case OP_MULHI64: r0 = mulhi64(a, b); break;
        case OP_SHOUP_ACC: r0 = shoup_acc(a, b, c, d, 0 - e); break;
        case OP_CSUB_N: r0 = csub_n(a, b, 0 - b); break;
        case OP_SUB64: r0 = sub64(a, b); break;
        case OP_TWICE_PLUS: r0 = twice_plus(a, b); break;
        case OP_FWD_BFLY: fwd_bfly(a, b, c, d, m); r0 = a; r1 = b; break;
        case OP_FWD_SKIP: fwd_bfly_relaxed<FwdStage::SKIP>(a, b, c, d, m); r0 = a; r1 = b; break;
        case OP_FWD_SUB8: fwd_bfly_relaxed<FwdStage::SUB8>(a, b, c, d, m); r0 = a; r1 = b; break;
        case OP_FWD_LAST: fwd_bfly_relaxed<FwdStage::LAST>(a, b, c, d, m); r0 = a; r1 = b; break;
        case OP_INV_BFLY: inv_bfly(a, b, c, d, m); r0 = a; r1 = b; break;
        case OP_INV_FINAL: inv_final_bfly(a, b, c, d, f, g, m); r0 = a; r1 = b; break;
        case OP_BARRETT_MUL:
        case OP_BARRETT_REDUCE: {
            Barrett bc;
            bc.q = e;
            bc.mu = mu;
            bc.s = s;
            bc.s_comp = s_comp;
            r0 = op == OP_BARRETT_MUL ? barrett_mul(a, b, bc) : barrett_reduce64(a, bc);
            break;
        }
        default: break;
    }
    out[2 * i] = r0;
    out[2 * i + 1] = r1;
}

}  // namespace

// Applies primitive `op` to `count` items: host arrays in (7 words per item)
// and out (2 words per item).  mu, s and s_comp are the Barrett constants of
// ops 11 and 12 (eltwise-mul-mod.cpp:22-30).  Returns 0 on success, -1 for
// bad arguments, the HIP error code otherwise.
extern "C" int axhel_arith_hooks_apply(int op, const uint64_t *in, uint64_t *out, uint64_t count, uint64_t mu,
                                       uint32_t s, uint32_t s_comp) {
    if (op < 0 || op >= OP_COUNT || !in || !out || count == 0 || count > (1ull << 26)) return -1;
    uint64_t *d_in = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(&d_in, count * 7 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&d_out, count * 2 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(d_in, in, count * 7 * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const unsigned threads = 256;
        const unsigned blocks = (unsigned)((count + threads - 1) / threads);
        hipLaunchKernelGGL(hooks_kernel, dim3(blocks), dim3(threads), 0, nullptr, op, d_in, d_out, count, mu, s,
                           s_comp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, count * 2 * sizeof(uint64_t), hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return e == hipSuccess ? 0 : (int)e;
}
