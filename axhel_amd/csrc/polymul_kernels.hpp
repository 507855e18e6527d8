// Fused RNS polynomial product, per limb:
//   out = InverseNTT( FMAMod( MulMod( NTT(a), NTT(b), 1 ), s, c, 1 ) )
// the call sequence of tests/public/test_polymul.cpp:13-25 with the FMA step
// of eltwise-fma-mod.cpp:37-77 in between (normalized transforms, factor 1).
//
// The forward transform's last pass and the inverse transform's first pass
// both act on the low LOG_G index bits of the same contiguous rows, so one
// kernel does, per row: the forward ROW stages of a and of b, the forward
// normalisation, MulMod, FMAMod with the addend row, and the inverse ROW
// stages, all in registers.  The forward pass ends with each thread holding
// E consecutive words and the inverse pass starts from exactly that layout,
// so no LDS exchange sits between them.  Every step is the reference's
// arithmetic on the same words, so the result is bit-identical to running
// the five calls one after another.  RLX = true runs the two forward
// transforms on the relaxed schedule of fwd_bfly_relaxed (every modulus
// below 2^60): their words differ on the way, but the normalised NTT(a) and
// NTT(b) the product consumes are the unique residues, so the output is
// unchanged.
#pragma once

#include "ntt_kernels.hpp"

namespace axhel_dev {

struct PolyJobDev {
    const uint64_t *a1;  // a after the forward COL pass (or a itself for a single-pass degree)
    const uint64_t *b1;  // b likewise
    const uint64_t *c;   // addend, natural order; null = none
    uint64_t *out;       // written here; the inverse COL pass then runs in place
    const ulonglong2 *twf, *twi;
    uint64_t q;
    Barrett bc;
    uint64_t scalar;     // FMA multiplier (ReduceInputNative<1>: as given)
    uint64_t ninv, ninv_q, sroot, sroot_q;
};

struct PolyArgs {
    const PolyJobDev *jobs;
    uint32_t logn;
    uint32_t num_limbs;
    uint32_t groups_per_limb;
    uint32_t single;  // the degree fits one ROW pass: apply the final stage and normalise here
    uint32_t vec16;   // addends 16-byte aligned
};

template <int LOG_G, int LOG_E>
struct FusedLayout {
    // the forward pass's last-chunk layout equals the inverse pass's first-chunk layout
    static constexpr bool ok() {
        using F = Geom<LOG_G, LOG_E, true>;
        using I = Geom<LOG_G, LOG_E, false>;
        for (int r = 0; r < (1 << LOG_E); ++r)
            if (F::kr(F::NC - 1, r) != I::kr(0, r)) return false;
        return F::LO(F::NC - 1) == 0 && I::LO(0) == 0;
    }
};

template <int LOG_G, int LOG_E, int GPW, bool RLX = false>
__global__ void __launch_bounds__(GPW << (LOG_G - LOG_E))
polymul_row_kernel(PolyArgs a) {
    static_assert(FusedLayout<LOG_G, LOG_E>::ok(), "forward exit layout must be the inverse entry layout");
    using Gf = Geom<LOG_G, LOG_E, true>;
    using Gi = Geom<LOG_G, LOG_E, false>;
    constexpr int G = 1 << LOG_G;
    constexpr int E = 1 << LOG_E;
    constexpr int T = G / E;
    extern __shared__ uint64_t lds[];

    uint32_t tile = blockIdx.x;
    {
        const uint32_t nb = gridDim.x;
        if ((nb & 7u) == 0) tile = (tile & 7u) * (nb >> 3) + (tile >> 3);
    }
    const uint32_t t = threadIdx.x;
    const uint32_t tid = t % T, g = t / T;
    const uint32_t gid = tile * GPW + g;
    const uint32_t lgpl = a.logn - LOG_G;  // groups_per_limb = 2^lgpl
    const uint32_t limb = (tile * GPW) >> lgpl;  // uniform (groups_per_limb % GPW == 0)
    const uint32_t o = gid - (limb << lgpl);
    const PolyJobDev P = a.jobs[limb];
    const ModConsts m = mod_consts(P.q);
    const uint32_t n = 1u << a.logn;
    const uint32_t row = o << LOG_G;

    NttJobDev fin{};
    fin.ninv = P.ninv;
    fin.ninv_q = P.ninv_q;
    fin.sroot = P.sroot;
    fin.sroot_q = P.sroot_q;

    uint64_t va[E], vb[E];
    const uint32_t k0 = Gf::kt(0, tid);
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        va[r] = ld_g(P.a1 + (row | (k0 + Gf::kr(0, r))));
    });
    const TwGlobal<LOG_G, true, false> twf{P.twf, n, 0, o};
    const TwGlobal<LOG_G, false, false> twi{P.twi, n, 0, o};
    // every LDS word a row touches is its own (Lds::idx2(g, ...)): when a row
    // lies inside one wave, its exchanges synchronize that wave alone
    constexpr bool WS = row_group_in_wave<LOG_G, LOG_E>();
    run_chunks<LOG_G, LOG_E, GPW, true, false, RLX, WS>(va, lds, g, tid, m, twf, false, fin);
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        vb[r] = ld_g(P.b1 + (row | (k0 + Gf::kr(0, r))));
    });
    group_sync<WS>();  // the LDS tile is reused by b's pass
    run_chunks<LOG_G, LOG_E, GPW, true, false, RLX, WS>(vb, lds, g, tid, m, twf, false, fin);

    // thread holds words row + kl + r, r in [0, E)
    const uint32_t kl = Gf::kt(Gf::NC - 1, tid);
    uint64_t cv[E];
    if (P.c) {
        if (a.vec16) {
            static_for<0, E / 2>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const ulonglong2 x2 = ld_nt2(P.c + (row | kl) + 2 * i);
                cv[2 * i] = x2.x;
                cv[2 * i + 1] = x2.y;
            });
        } else {
            static_for<0, E>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                cv[r] = ld_g(P.c + (row | kl) + r);
            });
        }
    }
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        // NTT normalisation (ntt.cpp:79-91), then MulMod and FMAMod at factor 1
        const uint64_t x = csub_n(csub_n(va[r], m.q2, m.nq2), m.q, m.nq);
        const uint64_t y = csub_n(csub_n(vb[r], m.q2, m.nq2), m.q, m.nq);
        const uint64_t prod = barrett_mul(x, y, P.bc);
        uint64_t d = barrett_mul(prod, P.scalar, P.bc);
        if (P.c) {
            d += cv[r];
            if (d >= m.q) d -= m.q;
        }
        va[r] = d;
    });
    run_chunks<LOG_G, LOG_E, GPW, false, false, false, WS>(va, lds, g, tid, m, twi, a.single != 0, fin);
    if (a.single) {
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            va[r] = csub_n(va[r], m.q, m.nq);
        });
    }
    const uint32_t ke = Gi::kt(Gi::NC - 1, tid);
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        st_g(P.out + (row | (ke + Gi::kr(Gi::NC - 1, r))), va[r]);
    });
}

// The fused ROW kernel over stacks of jobs sharing one plan (two-pass
// degrees): workgroup (stack s, row o) runs the fused steps on row o of the
// GPW jobs jobs[s*GPW ..]; empty slots have out == nullptr, slot 0 is always
// a job.  Row o's forward and inverse twiddles are staged in LDS once for all
// GPW rows (each row needs its own 2 x (2^LOG_G - 1) pairs, which the
// per-limb kernel reads from L2).  Same arithmetic as polymul_row_kernel.
template <int LOG_G, int LOG_E, int GPW, bool RLX = false>
__global__ void __launch_bounds__(GPW << (LOG_G - LOG_E))
polymul_row_stack_kernel(PolyArgs a) {
    static_assert(FusedLayout<LOG_G, LOG_E>::ok(), "forward exit layout must be the inverse entry layout");
    using Gf = Geom<LOG_G, LOG_E, true>;
    using Gi = Geom<LOG_G, LOG_E, false>;
    using L = Lds<LOG_G, LOG_E, GPW, false>;
    constexpr int G = 1 << LOG_G;
    constexpr int E = 1 << LOG_E;
    constexpr int T = G / E;
    constexpr int NT = GPW * T;
    extern __shared__ uint64_t lds[];

    uint32_t tile = blockIdx.x;
    {
        const uint32_t nb = gridDim.x;
        if ((nb & 7u) == 0) tile = (tile & 7u) * (nb >> 3) + (tile >> 3);
    }
    const uint32_t t = threadIdx.x;
    const uint32_t tid = t % T, g = t / T;
    const uint32_t lgpl = a.logn - LOG_G;  // groups_per_limb = 2^lgpl
    const uint32_t st = tile >> lgpl;
    const uint32_t o = tile & ((1u << lgpl) - 1u);
    const PolyJobDev P0 = a.jobs[st * GPW];  // the stack's plan
    const PolyJobDev *Pg = a.jobs + st * GPW + g;
    uint64_t *out = Pg->out;
    const bool valid = out != nullptr;
    const uint64_t *a1 = Pg->a1, *b1 = Pg->b1, *c = Pg->c;
    const uint64_t scalar = Pg->scalar;
    const ModConsts m = mod_consts(P0.q);
    const uint32_t n = 1u << a.logn;
    const uint64_t row = (uint64_t)o << LOG_G;

    ulonglong2 *twf_lds = reinterpret_cast<ulonglong2 *>(lds + L::WORDS);
    ulonglong2 *twi_lds = twf_lds + (G - 1);
    TwStage<LOG_G, NT> sf, si;
    sf.template fetch_row<true, LOG_E>(P0.twf, n, o, t);
    si.template fetch_row<false, LOG_E>(P0.twi, n, o, t);

    NttJobDev fin{};  // no final stage here: it belongs to the inverse COL pass

    uint64_t va[E], vb[E];
    const uint32_t k0 = Gf::kt(0, tid);
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        va[r] = valid ? ld_g(a1 + row + k0 + Gf::kr(0, r)) : 0;
    });
    sf.store(twf_lds, t);
    si.store(twi_lds, t);
    __syncthreads();  // publishes the twiddles
    // one row o and one plan per workgroup: the top chunk's entries are uniform
    const TwTopChunkUniform<top_chunk<LOG_G, LOG_E, true>(), TwGlobal<LOG_G, true, false>, TwLdsT<LOG_G, LOG_E, true>>
        twf{{P0.twf, n, 0u, o}, {twf_lds}};
    const TwTopChunkUniform<top_chunk<LOG_G, LOG_E, false>(), TwGlobal<LOG_G, false, false>,
                            TwLdsT<LOG_G, LOG_E, false>>
        twi{{P0.twi, n, 0u, o}, {twi_lds}};
    // the staged twiddles are shared (published above); every other LDS word
    // a row touches is its own, so a row inside one wave synchronizes alone
    constexpr bool WS = row_group_in_wave<LOG_G, LOG_E>();
    run_chunks<LOG_G, LOG_E, GPW, true, false, RLX, WS>(va, lds, g, tid, m, twf, false, fin);
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        vb[r] = valid ? ld_g(b1 + row + k0 + Gf::kr(0, r)) : 0;
    });
    group_sync<WS>();  // the LDS tile is reused by b's pass
    run_chunks<LOG_G, LOG_E, GPW, true, false, RLX, WS>(vb, lds, g, tid, m, twf, false, fin);

    const uint32_t kl = Gf::kt(Gf::NC - 1, tid);
    uint64_t cv[E];
    if (c && valid) {
        if (a.vec16) {
            static_for<0, E / 2>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const ulonglong2 x2 = ld_nt2(c + row + kl + 2 * i);
                cv[2 * i] = x2.x;
                cv[2 * i + 1] = x2.y;
            });
        } else {
            static_for<0, E>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                cv[r] = ld_g(c + row + kl + r);
            });
        }
    }
    static_for<0, E>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        // NTT normalisation (ntt.cpp:79-91), then MulMod and FMAMod at factor 1
        const uint64_t x = csub_n(csub_n(va[r], m.q2, m.nq2), m.q, m.nq);
        const uint64_t y = csub_n(csub_n(vb[r], m.q2, m.nq2), m.q, m.nq);
        const uint64_t prod = barrett_mul(x, y, P0.bc);
        uint64_t d = barrett_mul(prod, scalar, P0.bc);
        if (c && valid) {
            d += cv[r];
            if (d >= m.q) d -= m.q;
        }
        va[r] = d;
    });
    run_chunks<LOG_G, LOG_E, GPW, false, false, false, WS>(va, lds, g, tid, m, twi, false, fin);
    if (valid) {
        const uint32_t ke = Gi::kt(Gi::NC - 1, tid);
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            st_g(out + row + ke + Gi::kr(Gi::NC - 1, r), va[r]);
        });
    }
}

}  // namespace axhel_dev
