/*
 * axhel_oracle.c -- CPU restatement of the AXHEL Native path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it. The product (axhel_amd/) never links, loads or calls it.
 *
 * Every function restates one piece of the reference, statement by statement,
 * so that lazy outputs come out with the same bit patterns.  Citations are
 * file:line in the reference tree (massidonati/AXHEL).
 *
 * Parity pins (what checks each function beyond this restatement):
 *   tables            example/example.cpp:315-345 golden N=8 tables; closed form
 *   normalized NTT    tests/public/test_ntt.cpp:32-46 KAT; O(N^2) DFT closed form
 *   lazy NTT          congruence with the closed form at every stage, the lazy
 *                     ranges [0,4q)/[0,2q), the Shoup representative from its
 *                     big-integer definition, and the lazy words themselves
 *                     against Harvey 2014 Algorithm 4 and its Gentleman-Sande
 *                     counterpart restated over Python integers (twiddles
 *                     from psi, boundary words 0, q, 2q-1, 2q, 4q-1); the last
 *                     inverse stage's representative follows the reference's
 *                     fused scaling (ntt-native.cpp:101-141), restated there
 *                     too.  No reference-held vector exists for lazy bit
 *                     patterns.
 *   add/sub/mul/fma   closed form (a op b) mod q, exhaustive small q, example KATs
 *   reduce            closed form; Barrett mode KAT example/example.cpp:274-285
 *   mt19937_64        the C++ standard's 10000th-output check value
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>

typedef unsigned __int128 u128;

/* ---- L1 primitives (axhel/include/axhel/number-theory/uint-arith.hpp) ---- */

/* uint-arith.hpp:25-27 */
static uint64_t bit_width(uint64_t x) { return x ? 64u - (uint64_t)__builtin_clzll(x) : 0; }

/* uint-arith.hpp:40-44 */
static uint64_t mul_hi(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) >> 64); }

/* uint-arith.hpp:48-52: value*mult - hi(value*quot)*q, all mod 2^64 */
static uint64_t shoup_lazy(uint64_t v, uint64_t w, uint64_t wq, uint64_t q) {
    uint64_t qhat = mul_hi(v, wq);
    return v * w - qhat * q;
}

/* uint-arith.hpp:56-60 */
uint64_t oracle_shoup_quotient(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

/* modular-reduction.hpp:22-25 */
static uint64_t csub(uint64_t x, uint64_t b) { return x >= b ? x - b : x; }

/* modular-reduction.hpp:48-67 (ReduceInputNative<F>) */
static uint64_t reduce_input(uint64_t x, uint64_t q, int f) {
    if (f == 8) x = csub(x, 4 * q);
    if (f >= 4) x = csub(x, 2 * q);
    if (f >= 2) x = csub(x, q);
    return x;
}

uint64_t oracle_shoup_lazy(uint64_t v, uint64_t w, uint64_t wq, uint64_t q) { return shoup_lazy(v, w, wq, q); }

/* ---- Tables (axhel/ntt/ntt-tables.cpp:17-140) ---- */

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

/* ntt-tables.cpp:24-41 */
static uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        e >>= 1;
        if (e) b = mulmod(b, b, q);
    }
    return r;
}

/* ntt-tables.cpp:54-65: increment a bit-reversed counter over log2(n) bits */
static size_t brv_next(size_t cur, size_t n) {
    size_t bit = n >> 1;
    while (cur & bit) { cur ^= bit; bit >>= 1; }
    return cur ^ bit;
}

/*
 * ntt-tables.cpp:81-140.  fwd/inv are AoS {operand, quotient} arrays of n
 * entries (2n words); ninv gets {n^-1, quotient}.  Returns 0, or -1 when the
 * scratch allocation fails.
 */
int oracle_gen_tables(uint64_t n, uint64_t q, uint64_t psi, uint64_t *fwd, uint64_t *inv, uint64_t *ninv) {
    uint64_t *seq = (uint64_t *)malloc(n * sizeof(uint64_t));
    size_t r, k;
    if (!seq) return -1;
    /* ntt-tables.cpp:68-76: seq[k] = psi^k */
    seq[0] = 1;
    for (k = 1; k < n; ++k) seq[k] = mulmod(seq[k - 1], psi, q);
    /* ntt-tables.cpp:95-106: root_powers[i] = psi^brv(i) */
    r = 0;
    for (k = 0; k < n; ++k) {
        fwd[2 * k] = seq[r];
        fwd[2 * k + 1] = oracle_shoup_quotient(seq[r], q);
        if (k + 1 < n) r = brv_next(r, n);
    }
    /* ntt-tables.cpp:108-129: inv_root_powers[0] = {0,0};
       inv_root_powers[i] = psi^-(brv(i-1)+1).  brv(i-1) <= n-2 for i <= n-1,
       so the exponent stays inside seq. */
    {
        uint64_t ipsi = powmod(psi, q - 2, q);
        seq[0] = 1;
        for (k = 1; k < n; ++k) seq[k] = mulmod(seq[k - 1], ipsi, q);
    }
    inv[0] = 0;
    inv[1] = 0;
    r = 0;
    for (k = 1; k < n; ++k) {
        uint64_t w = seq[r + 1];
        inv[2 * k] = w;
        inv[2 * k + 1] = oracle_shoup_quotient(w, q);
        if (k + 1 < n) r = brv_next(r, n);
    }
    /* ntt-tables.cpp:131-136 */
    {
        uint64_t ni = powmod(n % q, q - 2, q);
        ninv[0] = ni;
        ninv[1] = oracle_shoup_quotient(ni, q);
    }
    free(seq);
    return 0;
}

/* ---- Forward lazy NTT (ntt-native.cpp:19-55, ntt-internal.hpp:20-31) ---- */

/* Runs the Cooley-Tukey stages m = 1, 2, ... while m < 2^stop_log_m
   (stop_log_m = log2(n) runs them all).  Exposed for per-stage checks. */
void oracle_fwd_lazy_stages(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *fwd, unsigned stop_log_m) {
    const uint64_t q2 = q << 1;
    const uint64_t m_stop = stop_log_m >= 63 ? n : ((uint64_t)1 << stop_log_m);
    uint64_t m = 1, gap = n >> 1;
    while (m < n && m < m_stop) {
        for (uint64_t i = 0; i < m; ++i) {
            const uint64_t w = fwd[2 * (m + i)], wq = fwd[2 * (m + i) + 1];
            uint64_t *x = a + i * (gap << 1);
            uint64_t *y = x + gap;
            for (uint64_t j = 0; j < gap; ++j) {
                uint64_t xh = csub(x[j], q2);                 /* [0,4q) -> [0,2q) */
                uint64_t t = shoup_lazy(y[j], w, wq, q);      /* [0,2q) */
                x[j] = xh + t;
                y[j] = xh + q2 - t;
            }
        }
        m <<= 1;
        gap >>= 1;
    }
}

void oracle_fwd_lazy(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *fwd) {
    oracle_fwd_lazy_stages(a, n, q, fwd, 64);
}

/* ---- Inverse lazy NTT (ntt-native.cpp:61-142, ntt-internal.hpp:37-49) ---- */

/* Runs the first `stages_to_run` Gentleman-Sande stages (m = n/2, n/4, ...,
   2; log2(n)-1 of them in all); final_stage != 0 then also runs the scaled
   last stage (m = 1) once every earlier stage has run. */
void oracle_inv_lazy_stages(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *inv, const uint64_t *ninv,
                            unsigned stages_to_run, int final_stage) {
    const uint64_t q2 = q << 1;
    uint64_t idx = 1, m = n >> 1, gap = 1;
    unsigned done = 0;
    while (m > 1 && done < stages_to_run) {
        for (uint64_t i = 0; i < m; ++i) {
            const uint64_t w = inv[2 * idx], wq = inv[2 * idx + 1];
            ++idx;
            uint64_t *x = a + i * (gap << 1);
            uint64_t *y = x + gap;
            for (uint64_t j = 0; j < gap; ++j) {
                uint64_t u = x[j], v = y[j];
                x[j] = csub(u + v, q2);
                y[j] = shoup_lazy(u + q2 - v, w, wq, q);
            }
        }
        m >>= 1;
        gap <<= 1;
        ++done;
    }
    if (!final_stage || m != 1) return;
    /* ntt-native.cpp:101-141 */
    {
        const uint64_t wl = inv[2 * idx];
        uint64_t s = shoup_lazy(wl, ninv[0], ninv[1], q);
        s = csub(s, q);
        const uint64_t sq = oracle_shoup_quotient(s, q);
        uint64_t *x = a, *y = a + gap;
        for (uint64_t j = 0; j < gap; ++j) {
            uint64_t xh = csub(x[j], q2);
            uint64_t yv = y[j];
            uint64_t sum = xh + yv;
            uint64_t dif = xh + q2 - yv;
            x[j] = shoup_lazy(sum, ninv[0], ninv[1], q);
            y[j] = shoup_lazy(dif, s, sq, q);
        }
    }
}

void oracle_inv_lazy(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *inv, const uint64_t *ninv) {
    oracle_inv_lazy_stages(a, n, q, inv, ninv, 64, 1);
}

/* ---- Element-wise kernels (native bodies in axhel/eltwise/eltwise-*-mod.cpp) ---- */

/* eltwise-add-mod.cpp:15-30 */
void oracle_add_vv(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t n, uint64_t q) {
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t s = a[i] + b[i];
        s -= (uint64_t)(s >= q) * q;
        r[i] = s;
    }
}
/* eltwise-add-mod.cpp:33-47 */
void oracle_add_vs(uint64_t *r, const uint64_t *a, uint64_t b, uint64_t n, uint64_t q) {
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t s = a[i] + b;
        s -= (uint64_t)(s >= q) * q;
        r[i] = s;
    }
}
/* eltwise-sub-mod.cpp:15-30 */
void oracle_sub_vv(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t n, uint64_t q) {
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x = a[i], y = b[i];
        uint64_t d = x - y;
        d += (uint64_t)(x < y) * q;
        r[i] = d;
    }
}
/* eltwise-sub-mod.cpp:33-47 */
void oracle_sub_vs(uint64_t *r, const uint64_t *a, uint64_t b, uint64_t n, uint64_t q) {
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x = a[i];
        uint64_t d = x - b;
        d += (uint64_t)(x < b) * q;
        r[i] = d;
    }
}

/* Barrett constants, eltwise-mul-mod.cpp:22-30 and multiply-factor.cpp:11-19:
   n = BitWidth(q), s = n - 2, mu = floor((2^(n-2) << 64) / q). */
static void barrett_consts(uint64_t q, uint64_t *s, uint64_t *mu) {
    uint64_t nb = bit_width(q);
    *s = nb - 2;
    *mu = (uint64_t)((((u128)1 << (nb - 2)) << 64) / q);
}

/* The 128-bit right shift (hi:lo) >> s as written at eltwise-mul-mod.cpp:37.
   With s == 0 the reference shifts hi left by 64, which AArch64 and x86-64
   register shifts both take modulo 64 (a shift by 0); the `& 63` spells that. */
static uint64_t shr128(uint64_t hi, uint64_t lo, uint64_t s) {
    return (lo >> s) | (hi << ((64 - s) & 63));
}

/* eltwise-mul-mod.cpp:19-63 + dispatch :81-94 (factors outside {1,2,4}: no-op) */
void oracle_mul(uint64_t *r, const uint64_t *a, const uint64_t *b, uint64_t n, uint64_t q, uint64_t f) {
    if (f != 1 && f != 2 && f != 4) return;
    uint64_t s, mu;
    barrett_consts(q, &s, &mu);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x = reduce_input(a[i], q, (int)f);
        uint64_t y = reduce_input(b[i], q, (int)f);
        u128 p = (u128)x * y;
        uint64_t c1 = shr128((uint64_t)(p >> 64), (uint64_t)p, s);
        uint64_t qh = mul_hi(c1, mu);
        uint64_t z = (uint64_t)p - qh * q;
        r[i] = reduce_input(z, q, 4);
    }
}

/* eltwise-fma-mod.cpp:19-78 + dispatch :96-112 (factors outside {1,2,4,8}: no-op) */
void oracle_fma(uint64_t *r, const uint64_t *a, uint64_t b, const uint64_t *c, uint64_t n, uint64_t q, uint64_t f) {
    if (f != 1 && f != 2 && f != 4 && f != 8) return;
    uint64_t s, mu;
    barrett_consts(q, &s, &mu);
    const uint64_t y = reduce_input(b, q, (int)f);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x = reduce_input(a[i], q, (int)f);
        u128 p = (u128)x * y;
        uint64_t c1 = shr128((uint64_t)(p >> 64), (uint64_t)p, s);
        uint64_t qh = mul_hi(c1, mu);
        uint64_t z = (uint64_t)p - qh * q;
        uint64_t v = reduce_input(z, q, 4);
        if (c) {
            v += reduce_input(c[i], q, (int)f);
            if (v >= q) v -= q;
        }
        r[i] = v;
    }
}

/* eltwise-reduce-mod.cpp:19-128: ordered branch list; anything else is a no-op */
void oracle_reduce(uint64_t *r, const uint64_t *a, uint64_t n, uint64_t q, uint64_t fin, uint64_t fout) {
    uint64_t i;
    if (fin == 1 || (fout == 2 && fin <= 2)) {
        for (i = 0; i < n; ++i) r[i] = a[i];
        return;
    }
    if (fin == 2 && fout == 1) {
        for (i = 0; i < n; ++i) r[i] = csub(a[i], q);
        return;
    }
    if (fin == 4 && fout == 1) {
        for (i = 0; i < n; ++i) r[i] = csub(csub(a[i], 2 * q), q);
        return;
    }
    if (fin == 4 && fout == 2) {
        for (i = 0; i < n; ++i) r[i] = csub(a[i], 2 * q);
        return;
    }
    if (fin == q) {
        uint64_t s, mu;
        barrett_consts(q, &s, &mu);
        for (i = 0; i < n; ++i) {
            uint64_t x = a[i];
            uint64_t qh = mul_hi(x >> s, mu);
            uint64_t z = x - qh * q;
            r[i] = reduce_input(z, q, 4);
        }
    }
}

/* ---- NTT dispatch (axhel/ntt/ntt.cpp:79-118, 173-204) ---- */

void oracle_compute_forward(uint64_t *res, const uint64_t *op, uint64_t n, uint64_t q, const uint64_t *fwd,
                            uint64_t in_f, uint64_t out_f) {
    (void)in_f; /* ignored by the reference, ntt.cpp:173-186 */
    if (res != op) memmove(res, op, n * sizeof(uint64_t));
    oracle_fwd_lazy(res, n, q, fwd);
    if (out_f != 4) oracle_reduce(res, res, n, q, 4, 1);
}

void oracle_compute_inverse(uint64_t *res, const uint64_t *op, uint64_t n, uint64_t q, const uint64_t *inv,
                            const uint64_t *ninv, uint64_t in_f, uint64_t out_f) {
    (void)in_f; /* ignored, ntt.cpp:192-204 */
    if (res != op) memmove(res, op, n * sizeof(uint64_t));
    oracle_inv_lazy(res, n, q, inv, ninv);
    if (out_f != 2) oracle_reduce(res, res, n, q, 2, 1);
}

/* ---- Test-parameter generators (tests/common/parameters.hpp) ---- */

static uint64_t powmod_t(uint64_t a, uint64_t e, uint64_t q) {
    /* tests/common/reference.hpp:29-37 */
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

/* parameters.hpp:29-51, deterministic Miller-Rabin for 64-bit n */
int oracle_is_prime(uint64_t n) {
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    unsigned i, r, s = 0;
    uint64_t d;
    if (n < 2) return 0;
    for (i = 0; i < 12; ++i)
        if (n % small[i] == 0) return n == small[i];
    d = n - 1;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (i = 0; i < 7; ++i) {
        uint64_t a = bases[i], x;
        int witness = 1;
        if (a % n == 0) continue;
        x = powmod_t(a % n, d, n);
        if (x == 1 || x == n - 1) continue;
        for (r = 1; r < s; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) { witness = 0; break; }
        }
        if (witness) return 0;
    }
    return 1;
}

/* parameters.hpp:58-70: largest prime q < 2^bits with q == 1 mod 2*degree (0 if none) */
uint64_t oracle_find_ntt_prime(uint64_t degree, unsigned bits) {
    const uint64_t step = 2 * degree;
    const uint64_t hi = ((uint64_t)1 << bits) - 1;
    const uint64_t lo = (uint64_t)1 << (bits - 1);
    uint64_t q = hi - ((hi - 1) % step);
    while (q >= lo && q > step) {
        if (oracle_is_prime(q)) return q;
        if (q < step) break;
        q -= step;
    }
    return 0;
}

/* parameters.hpp:110-122 */
uint64_t oracle_find_primitive_2n_root(uint64_t degree, uint64_t q) {
    const uint64_t order = 2 * degree;
    const uint64_t e = (q - 1) / order;
    uint64_t a;
    for (a = 2; a < 100000; ++a) {
        uint64_t psi = powmod_t(a, e, q);
        if (psi != 1 && powmod_t(psi, degree, q) == q - 1 && powmod_t(psi, order, q) == 1) return psi;
    }
    return 0;
}

/* mt19937_64 (Matsumoto & Nishimura 2004; parameters as fixed by the C++
   standard [rand.predef]).  The state lives in the caller's 313-word buffer:
   words 0..311 = state, word 312 = index. */
#define MT_N 312
#define MT_M 156
void oracle_mt_seed(uint64_t *st, uint64_t seed) {
    unsigned i;
    st[0] = seed;
    for (i = 1; i < MT_N; ++i) st[i] = 6364136223846793005ULL * (st[i - 1] ^ (st[i - 1] >> 62)) + i;
    st[MT_N] = MT_N;
}

uint64_t oracle_mt_next(uint64_t *st) {
    uint64_t x;
    if (st[MT_N] >= MT_N) {
        unsigned i;
        for (i = 0; i < MT_N; ++i) {
            uint64_t y = (st[i] & 0xFFFFFFFF80000000ULL) | (st[(i + 1) % MT_N] & 0x7FFFFFFFULL);
            uint64_t v = st[(i + MT_M) % MT_N] ^ (y >> 1);
            if (y & 1) v ^= 0xB5026F5AA96619E9ULL;
            st[i] = v;
        }
        st[MT_N] = 0;
    }
    x = st[st[MT_N]++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= x >> 43;
    return x;
}

/* parameters.hpp:253-270: rejection-sampled uniform values in [0, bound) */
void oracle_random_vector(uint64_t *st, uint64_t bound, uint64_t *out, uint64_t n) {
    const uint64_t threshold = (uint64_t)(0 - bound) % bound;
    uint64_t i;
    for (i = 0; i < n; ++i) {
        for (;;) {
            uint64_t x = oracle_mt_next(st);
            if (x >= threshold) { out[i] = x % bound; break; }
        }
    }
}
