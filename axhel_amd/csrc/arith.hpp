// Exact 64-bit modular primitives for gfx950 device code.
//
// These are the device forms of the scalar routines the reference's Native
// backend uses (axhel/include/axhel/number-theory/uint-arith.hpp:19-60 and
// modular-reduction.hpp:22-67).  Each returns the same 64-bit value as the
// reference for every input, including out-of-range ones: products are exact
// hi64/lo64 parts, never approximations, because lazy NTT outputs depend on
// the exact Shoup quotient.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axhel_dev {

// x >= b ? x - b : x  (modular-reduction.hpp:22-25).
__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t b) {
    const uint64_t d = x - b;
    return x >= b ? d : x;
}

// ReduceInputNative<F> (modular-reduction.hpp:48-67): cascaded conditional
// subtractions 4q, 2q, q for F = 8; 2q, q for F = 4; q for F = 2.
template <int F>
__device__ __forceinline__ uint64_t reduce_input(uint64_t x, uint64_t q) {
    if constexpr (F == 8) x = csub(x, q << 2);
    if constexpr (F >= 4) x = csub(x, q << 1);
    if constexpr (F >= 2) x = csub(x, q);
    return x;
}

// Shoup lazy product y*w - hi64(y*w')*q mod 2^64, in [0, 2q) when w < q and
// w' = floor(w * 2^64 / q) (uint-arith.hpp:48-52).
__device__ __forceinline__ uint64_t shoup(uint64_t y, uint64_t w, uint64_t wq, uint64_t q) {
    const uint64_t qhat = __umul64hi(y, wq);
    return y * w - qhat * q;
}

// gfx950 VALU forms.  Two inline-assembly pieces expose what the compiler
// does not: the carry-out of v_mad_u64_u32 and a select driven by the carry
// of a 64-bit add.  A VALU that writes an SGPR (a carry or lane mask) needs
// two wait states before a VALU reads that SGPR; the compiler does not see
// inside these blocks, so each block carries its own s_nop and keeps every
// SGPR it writes to itself.  All operands are VGPRs, so the forms hold for
// per-lane as well as uniform values.

// u + carry * 2^64 = a * b + t; carry returned as 0 or 1.
__device__ __forceinline__ uint64_t mad_u64_carry(uint32_t a, uint32_t b, uint64_t t, uint32_t &carry) {
    uint64_t u, c;
    asm("v_mad_u64_u32 %0, %1, %3, %4, %5\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %2, 0, 1, %1"
        : "=v"(u), "=s"(c), "=v"(carry)
        : "v"(a), "v"(b), "v"(t));
    return u;
}

__device__ __forceinline__ uint32_t add3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Exact hi64(a * b) in six VALU instructions: the middle column
// a0*b1 + a1*b0 + hi(a0*b0) can reach 2^65, and its carry is taken from the
// second v_mad_u64_u32 instead of being rebuilt from zero-extended halves.
__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t t = (uint64_t)a1 * b0 + __umulhi(a0, b0);  // < 2^64
    uint32_t carry;
    const uint64_t u = mad_u64_carry(a0, b1, t, carry);
    return (uint64_t)a1 * b1 + ((u >> 32) | ((uint64_t)carry << 32));
}

// acc + shoup(y, w, wq, q) mod 2^64, spelled for the gfx950 VALU.  With
// nq = 2^64 - q, y*w - qhat*q == y*w + qhat*nq (mod 2^64): the two low
// products chain through one v_mad_u64_u32 accumulator that starts at acc,
// and the four cross terms (two per v_mul_lo/v_mad pair) land in the high
// word with one 3-input add.  acc = 0 gives shoup() itself.
__device__ __forceinline__ uint64_t shoup_acc(uint64_t acc, uint64_t y, uint64_t w, uint64_t wq, uint64_t nq) {
    const uint64_t qhat = mulhi64(y, wq);
    const uint32_t yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
    const uint32_t wl = (uint32_t)w, wh = (uint32_t)(w >> 32);
    const uint32_t ql = (uint32_t)qhat, qh = (uint32_t)(qhat >> 32);
    const uint32_t nl = (uint32_t)nq, nh = (uint32_t)(nq >> 32);
    uint64_t p = (uint64_t)yl * wl + acc;
    p = (uint64_t)ql * nl + p;
    const uint32_t c1 = yl * wh + yh * wl;
    const uint32_t c2 = ql * nh + qh * nl;
    const uint32_t hi = add3_u32((uint32_t)(p >> 32), c1, c2);
    return (p & 0xffffffffull) | ((uint64_t)hi << 32);
}

__device__ __forceinline__ uint64_t shoup_n(uint64_t y, uint64_t w, uint64_t wq, uint64_t nq) {
    return shoup_acc(0, y, w, wq, nq);
}

// a - b mod 2^64 as one borrow chain.  Left to itself the compiler splits a
// difference whose subtrahend was assembled from two halves (as shoup_acc's
// result is) into several partial subtractions.
__device__ __forceinline__ uint64_t sub64(uint64_t a, uint64_t b) {
    uint32_t lo, hi;
    uint64_t c;
    asm("v_sub_co_u32_e64 %0, %2, %3, %5\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %1, %2, %4, %6, %2"
        : "=&v"(lo), "=v"(hi), "=&s"(c)
        : "v"((uint32_t)a), "v"((uint32_t)(a >> 32)), "v"((uint32_t)b), "v"((uint32_t)(b >> 32)));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// 2a + b in one v_lshl_add_u64 (the compiler otherwise shifts and adds).
__device__ __forceinline__ uint64_t twice_plus(uint64_t a, uint64_t b) {
    uint64_t r;
    asm("v_lshl_add_u64 %0, %1, 1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Per-limb constants of the butterflies: q, 2q, 4q, 8q and their negations
// mod 2^64 (4q and 8q serve the relaxed schedule, whose moduli are below 2^60).
struct ModConsts {
    uint64_t q, q2, nq, nq2, q4, nq4, q8, nq8;
};

__device__ __forceinline__ ModConsts mod_consts(uint64_t q) {
    ModConsts m;
    m.q = q;
    m.q2 = q << 1;
    m.nq = 0 - q;
    m.nq2 = 0 - m.q2;
    m.q4 = q << 2;
    m.nq4 = 0 - m.q4;
    m.q8 = q << 3;
    m.nq8 = 0 - m.q8;
    return m;
}

// csub(x, b) given nb = 2^64 - b (b > 0): x + nb carries out exactly when
// x >= b, and that carry selects between x + nb and x (four VALU
// instructions).
__device__ __forceinline__ uint64_t csub_n(uint64_t x, uint64_t b, uint64_t nb) {
    (void)b;
    uint32_t lo, hi;
    uint64_t c;
    asm("v_add_co_u32_e64 %0, %2, %3, %5\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %1, %2, %4, %6, %2\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %0, %3, %0, %2\n\t"
        "v_cndmask_b32_e64 %1, %4, %1, %2"
        : "=&v"(lo), "=&v"(hi), "=&s"(c)
        : "v"((uint32_t)x), "v"((uint32_t)(x >> 32)), "v"((uint32_t)nb), "v"((uint32_t)(nb >> 32)));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Harvey forward (Cooley-Tukey) butterfly, ntt-internal.hpp:20-31:
// x' = csub(x, 2q), t = shoup(y), x = x' + t, y = x' + 2q - t.  The sum
// comes out of the Shoup accumulator and the difference as
// (2x' + 2q) - (x' + t); both are below 4q, so the wrap-around arithmetic
// gives the reference's values exactly.
__device__ __forceinline__ void fwd_bfly(uint64_t &x, uint64_t &y, uint64_t w, uint64_t wq, const ModConsts &m) {
    const uint64_t xh = csub_n(x, m.q2, m.nq2);
    const uint64_t sum = shoup_acc(xh, y, w, wq, m.nq);
    y = sub64(twice_plus(xh, m.q2), sum);
    x = sum;
}

// Forward butterfly of the relaxed schedule, for transforms whose output is
// normalized to [0, q) and whose moduli are all below 2^60.  A normalized
// output is the unique canonical residue, so any congruent intermediate
// values give the reference's result bit for bit (for inputs in its domain
// [0, 4q), ntt-native.cpp:17); the schedule lets words grow to 16q and trades
// the per-stage bound for fewer conditional subtractions.  Every stage reads
// x' + t and x' + 2q - t with t = Shoup(y) in [0, 2q) for any 64-bit y, so a
// stage without subtraction raises the bound on its words by 2q:
//   SKIP: x' = x                        x < 14q -> outputs < 16q
//   SUB8: x' = CS(x, 8q)                x < 16q -> x' < 8q, outputs < 10q
//   LAST: x' = CS(CS(CS(x,8q),4q),2q)   x < 16q -> x' < 2q, outputs < 4q
// The passes place one SUB8 in every four stages (a first pass starts from
// inputs below 4q and skips its first six), which keeps every word below
// 16q < 2^64; LAST, the transform's bit 0, leaves the [0, 4q) range the
// normalization of the exact schedule expects (stage_kind in ntt_kernels.hpp).
enum class FwdStage { SKIP, SUB8, LAST };

template <FwdStage S>
__device__ __forceinline__ void fwd_bfly_relaxed(uint64_t &x, uint64_t &y, uint64_t w, uint64_t wq,
                                                 const ModConsts &m) {
    uint64_t xh = x;
    if constexpr (S != FwdStage::SKIP) xh = csub_n(xh, m.q8, m.nq8);
    if constexpr (S == FwdStage::LAST) xh = csub_n(csub_n(xh, m.q4, m.nq4), m.q2, m.nq2);
    const uint64_t sum = shoup_acc(xh, y, w, wq, m.nq);
    y = sub64(twice_plus(xh, m.q2), sum);
    x = sum;
}

// Harvey inverse (Gentleman-Sande) butterfly, ntt-internal.hpp:37-49.
__device__ __forceinline__ void inv_bfly(uint64_t &x, uint64_t &y, uint64_t w, uint64_t wq, const ModConsts &m) {
    const uint64_t u = x, v = y;
    x = csub_n(u + v, m.q2, m.nq2);
    y = shoup_n(u + m.q2 - v, w, wq, m.nq);
}

// Final inverse stage butterfly with n^-1 and the scaled root folded in,
// ntt-native.cpp:126-141.
__device__ __forceinline__ void inv_final_bfly(uint64_t &x, uint64_t &y, uint64_t ni, uint64_t niq, uint64_t s,
                                               uint64_t sq, const ModConsts &m) {
    const uint64_t xh = csub_n(x, m.q2, m.nq2);
    const uint64_t sum = xh + y;
    const uint64_t dif = xh + m.q2 - y;
    x = shoup_n(sum, ni, niq, m.nq);
    y = shoup_n(dif, s, sq, m.nq);
}

// Barrett constants of eltwise-mul-mod.cpp:22-30 (n = BitWidth(q),
// s = n - 2, mu = floor(2^(n-2) * 2^64 / q)), computed once on the host.
struct Barrett {
    uint64_t q;
    uint64_t mu;
    uint32_t s;
    uint32_t s_comp;  // (64 - s) & 63: the reference's hi << (64 - s) taken mod 64
};

// 128-bit (hi:lo) >> s as at eltwise-mul-mod.cpp:37.  For s == 0 the
// reference's `hi << 64` is a register shift, which both AArch64 and x86-64
// reduce modulo 64 (shift by 0); s_comp carries that reduction.
__device__ __forceinline__ uint64_t shr128(uint64_t hi, uint64_t lo, const Barrett &b) {
    return (lo >> b.s) | (hi << b.s_comp);
}

// Barrett product of x, y < q, eltwise-mul-mod.cpp:34-61: z in [0, 4q) then
// reduced to [0, q).
__device__ __forceinline__ uint64_t barrett_mul(uint64_t x, uint64_t y, const Barrett &b) {
    const uint64_t lo = x * y;
    const uint64_t hi = __umul64hi(x, y);
    const uint64_t c1 = shr128(hi, lo, b);
    const uint64_t qhat = __umul64hi(c1, b.mu);
    const uint64_t z = lo - qhat * b.q;
    return reduce_input<4>(z, b.q);
}

// Barrett reduction of an arbitrary 64-bit word, eltwise-reduce-mod.cpp:19-29.
__device__ __forceinline__ uint64_t barrett_reduce64(uint64_t x, const Barrett &b) {
    const uint64_t qhat = __umul64hi(x >> b.s, b.mu);
    const uint64_t z = x - qhat * b.q;
    return reduce_input<4>(z, b.q);
}

}  // namespace axhel_dev
