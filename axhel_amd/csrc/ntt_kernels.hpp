// Negacyclic Harvey NTT pass kernels for gfx950.
//
// A length-N transform (N = 2^logN, uint64 words) applies radix-2 Harvey
// butterflies over the bit positions of the coefficient index: the forward
// (Cooley-Tukey) transform from bit logN-1 down to bit 0, the inverse
// (Gentleman-Sande) from bit 0 up to bit logN-1 (ntt-native.cpp:31-54 and
// :72-141).  A butterfly on bit B pairs P and P + 2^B (bit B of P clear) and
// uses twiddle
//     forward:  root_powers[(N >> (B+1)) + (P >> (B+1))]        (= [m + i])
//     inverse:  inv_root_powers[N + 1 - (N >> B) + (P >> (B+1))] (= [N-2m+1+i])
// with m = N >> (B+1) blocks, i = P >> (B+1) the block index; the inverse's
// last bit (B = logN-1) is the scaled final stage of ntt-native.cpp:101-141.
//
// Every butterfly is the reference's radix-2 formula on the reference's pair
// with the reference's twiddle, and butterflies on one bit run only after all
// butterflies on the previous bit that feed them, so outputs are bit-identical
// to the scalar loop, lazy bits included.  Only the order of independent
// butterflies changes.  The one exception is chosen by the host for forward
// transforms with a normalized output and every q < 2^60 (RLX = true): the
// conditional subtractions follow the relaxed schedule of fwd_bfly_relaxed,
// whose intermediate words differ but whose normalized outputs, the unique
// residues in [0, q), are the reference's for inputs in [0, 4q).
//
// Mapping to the hardware.  A pass handles a contiguous range of LOG_G bit
// positions.  Fixing every other index bit leaves a "group" of G = 2^LOG_G
// words that the pass transforms independently:
//   ROW pass: bits [0, LOG_G) -> the group is G contiguous words;
//   COL pass: bits [logN-LOG_G, logN) -> the group is G words at stride
//             2^(logN-LOG_G) (a column of the N/G x G row-major matrix).
// Small N can run as one ROW pass (the whole limb sits in LDS; every
// configuration of N <= 2^11, and the batch configurations of 2^12 and
// 2^13); larger N, and the single-call shapes of 2^12 and 2^13, run a COL
// pass and a ROW pass (forward: COL then ROW; inverse: ROW then COL), two
// HBM round trips in all.
//
// Inside a group, T = G/E threads each hold E = 2^LOG_E words in registers.
// The pass's bits are cut into chunks of LOG_E bits (the last may be
// shorter); a chunk's butterflies all run in registers, and between chunks the
// group re-distributes through LDS so the next chunk's bits are register-
// local.  A workgroup holds GPW groups: COL groups are adjacent columns, so a
// wave's lanes walk adjacent addresses; ROW groups are adjacent rows.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "arith.hpp"

namespace axhel_dev {

// Device descriptor of one limb-transform.
struct NttJobDev {
    const uint64_t *src;    // operand (read by the first pass)
    uint64_t *dst;          // result (written by every pass; may equal src)
    const ulonglong2 *tw;   // forward or inverse table, {operand, quotient} AoS
    uint64_t q;
    uint64_t ninv, ninv_q;  // n^-1 and its Shoup quotient (inverse final stage)
    uint64_t sroot, sroot_q;  // scaled final root CS(Shoup(W_last, n^-1), q) and quotient
};

struct PassArgs {
    const NttJobDev *jobs;  // device job table; null -> inline_job (single limb)
    NttJobDev inline_job;
    uint32_t logn;
    uint32_t num_limbs;
    uint32_t groups_per_limb;  // N >> LOG_G (rows of a ROW pass, columns of a COL pass)
    uint32_t first_pass;       // read the job's src (else dst)
    uint32_t normalize;        // last pass: fold the [0,4q)->[0,q) / [0,2q)->[0,q) reduction in
    uint32_t final_inv;        // inverse pass holding bit logN-1 (scaled final stage)
    uint32_t vec16;            // every job's src and dst are 16-byte aligned
};

// The job's descriptor: from the device table, or the inline copy in the
// kernel arguments.  Fields are read one by one from a global-space pointer so
// the two sources stay separate loads (a select between their addresses would
// need generic pointers and a private copy of the kernel arguments).
__device__ __forceinline__ NttJobDev load_job(const PassArgs &a, uint32_t limb) {
    if (!a.jobs) return a.inline_job;
    typedef __attribute__((address_space(1))) const NttJobDev gjob;
    const gjob *p = (const gjob *)a.jobs + limb;
    NttJobDev j;
    j.src = p->src;
    j.dst = p->dst;
    j.tw = p->tw;
    j.q = p->q;
    j.ninv = p->ninv;
    j.ninv_q = p->ninv_q;
    j.sroot = p->sroot;
    j.sroot_q = p->sroot_q;
    return j;
}

__device__ __forceinline__ const uint64_t *job_src(const PassArgs &a, uint32_t limb) {
    if (!a.jobs) return a.first_pass ? a.inline_job.src : a.inline_job.dst;
    typedef __attribute__((address_space(1))) const NttJobDev gjob;
    const gjob *p = (const gjob *)a.jobs + limb;
    return a.first_pass ? p->src : p->dst;
}

__device__ __forceinline__ uint64_t *job_dst(const PassArgs &a, uint32_t limb) {
    if (!a.jobs) return a.inline_job.dst;
    typedef __attribute__((address_space(1))) const NttJobDev gjob;
    return ((const gjob *)a.jobs + limb)->dst;
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Geometry of chunk C of a pass with LOG_G bits and LOG_E bits per chunk.
template <int LOG_G, int LOG_E, bool FWD>
struct Geom {
    static constexpr int NC = (LOG_G + LOG_E - 1) / LOG_E;
    static constexpr int LAST_B = LOG_G - (NC - 1) * LOG_E;
    static constexpr int B(int c) { return c < NC - 1 ? LOG_E : LAST_B; }
    // Forward chunks walk the local bits from the top, inverse from the bottom.
    static constexpr int LO(int c) { return FWD ? LOG_G - (c * LOG_E + B(c)) : c * LOG_E; }
    // Inserts the B-bit gap at position LO into an (LOG_G - B)-bit value.
    static constexpr uint32_t spread(uint32_t o, int lo, int b) {
        return (o & ((1u << lo) - 1u)) | ((o >> lo) << (lo + b));
    }
    // Local index of register r (compile-time part) in chunk c.
    static constexpr uint32_t kr(int c, int r) {
        const int b = B(c), lo = LO(c);
        const uint32_t rlow = (uint32_t)r & ((1u << b) - 1u);
        const uint32_t rhi = (uint32_t)r >> b;
        return spread(rhi, lo, b) | (rlow << lo);
    }
    // Local index contribution of the in-group thread id in chunk c.
    __device__ static uint32_t kt(int c, uint32_t tid) {
        const int b = B(c), lo = LO(c);
        return spread(tid << (LOG_E - b), lo, b);
    }
    // Smallest register sharing r's twiddle on local bit p of chunk c (the
    // twiddle depends only on index bits above p).
    static constexpr int rep(int c, int r, int p) {
        const uint32_t key = kr(c, r) >> (p + 1);
        const int bit = 1 << (p - LO(c));
        for (int s = 0; s < (1 << LOG_E); ++s) {
            if (s & bit) continue;
            if ((kr(c, s) >> (p + 1)) == key) return s;
        }
        return r;
    }
};

template <int LOG_G, int LOG_E, int GPW, bool COL>
struct Lds {
    static constexpr int G = 1 << LOG_G;
    // ROW groups pad one word per E words so thread-contiguous chunks hit
    // distinct banks; COL tiles interleave groups so lanes (adjacent groups)
    // hit adjacent words.
    static constexpr int ROW_STRIDE = G + (G >> LOG_E);
    static constexpr int WORDS = COL ? GPW * G : GPW * ROW_STRIDE;
    // COL passes stage the pass's G - 1 twiddle pairs after the data tile
    // (pass_tile_words).
    static constexpr int TW_WORDS = COL ? 2 * (G - 1) : 0;
    __device__ static uint32_t idx(uint32_t g, uint32_t k) {
        if constexpr (COL) return k * GPW + g;
        else return g * ROW_STRIDE + k + (k >> LOG_E);
    }
    // idx(g, kt | kr) for a thread part kt and a register part kr with
    // disjoint bits (Geom's kt and kr always are): the padding then splits,
    // k + (k >> LOG_E) = [kt + (kt >> LOG_E)] + [kr + (kr >> LOG_E)], and a
    // compile-time kr folds into the LDS instruction's offset.
    __device__ static uint32_t idx2(uint32_t g, uint32_t kt, uint32_t kr) {
        if constexpr (COL) return (kt + kr) * GPW + g;
        else return g * ROW_STRIDE + (kt + (kt >> LOG_E)) + (kr + (kr >> LOG_E));
    }
};

// Global-memory accesses through address-space-1 pointers, so they issue as
// global_* (not flat_*) instructions and do not count against LDS waits.
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) uint64_t gword;
typedef __attribute__((address_space(1))) u64x2 gword2;

// Coefficient traffic streams through once per pass: it is loaded and stored
// non-temporally so it does not evict the twiddle tables from L2 (twiddles,
// read through ld_g2, stay cacheable).
__device__ __forceinline__ uint64_t ld_g(const uint64_t *p) { return __builtin_nontemporal_load((const gword *)p); }
__device__ __forceinline__ void st_g(uint64_t *p, uint64_t v) { __builtin_nontemporal_store(v, (gword *)p); }
// base + byte_off with a 32-bit offset: a uniform base can then stay in
// SGPRs (the instruction's scalar base) and the offset is one VGPR.
__device__ __forceinline__ uint64_t ld_gb(const uint64_t *base, uint32_t byte_off) {
    return __builtin_nontemporal_load((const gword *)((const char *)base + byte_off));
}
__device__ __forceinline__ void st_gb(uint64_t *base, uint32_t byte_off, uint64_t v) {
    __builtin_nontemporal_store(v, (gword *)((char *)base + byte_off));
}
__device__ __forceinline__ ulonglong2 ld_g2(const void *p) {
    const u64x2 x = *(const gword2 *)p;
    ulonglong2 r;
    r.x = x.x;
    r.y = x.y;
    return r;
}
__device__ __forceinline__ ulonglong2 ld_nt2(const void *p) {
    const u64x2 x = __builtin_nontemporal_load((const gword2 *)p);
    ulonglong2 r;
    r.x = x.x;
    r.y = x.y;
    return r;
}
__device__ __forceinline__ void st_g2(void *p, ulonglong2 v) {
    u64x2 x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, (gword2 *)p);
}

// Twiddles of a pass.  On local bit p (j = LOG_G - 1 - p) a pass uses 2^j
// consecutive table entries, from
//     forward:  (N >> (gb+1)) + (o << j)          gb = p + blo, the global bit
//     inverse:  N + 1 - (N >> gb) + (o << j)
// where the (o << j) term (the row) appears in ROW passes only; lp in
// [0, 2^j) selects the entry.  TwGlobal reads the table in HBM/L2; TwLds reads
// a copy staged in LDS, entry (j, lp) at
//     forward:  2^j - 1 + lp        inverse:  2^LOG_G - 2^(j+1) + lp
// (both orders cover [0, 2^LOG_G - 1) once).  In a COL pass both orders
// match the table's own: the copy is entries [base, base + 2^LOG_G - 1) with
// base = 1 (forward) or N + 1 - 2^LOG_G (inverse).
template <int LOG_G, bool FWD>
__device__ __forceinline__ constexpr uint32_t tw_lds_index(int j, uint32_t lp) {
    return (FWD ? (1u << j) - 1u : (1u << LOG_G) - (2u << j)) + lp;
}

// Every accessor takes the thread's part of the local index in the current
// chunk (Geom::kt) and the compile-time register part of lp; entry
// lp = (kt >> (P+1)) + pr.
template <int LOG_G, bool FWD, bool COL>
struct TwGlobal {
    const ulonglong2 *tw;
    uint32_t n, blo, o;
    template <int P>
    __device__ __forceinline__ ulonglong2 get(uint32_t kt, uint32_t pr) const {
        constexpr int j = LOG_G - 1 - P;
        const uint32_t gb = P + blo;
        uint32_t base = FWD ? (n >> (gb + 1)) : n + 1u - (n >> gb);
        if constexpr (!COL) base += o << j;
        return ld_g2(tw + base + (kt >> (P + 1)) + pr);
    }
};

template <int LOG_G, bool FWD>
struct TwLds {
    const ulonglong2 *t;
    template <int P>
    __device__ __forceinline__ ulonglong2 get(uint32_t kt, uint32_t pr) const {
        constexpr int j = LOG_G - 1 - P;
        return t[tw_lds_index<LOG_G, FWD>(j, (kt >> (P + 1)) + pr)];
    }
};

// Transposed LDS layout of a ROW pass's twiddles.  Stage p lies in chunk c
// (bits [LO, LO+B) in registers).  A chunk narrower than the radix (B <
// LOG_E, only ever the last) also keeps the thread's spare register bits,
// and Geom::kr places those that do not fit below LO just above the chunk,
// so the registers cover [LO, TOP) with TOP = LO + B + max(0, LOG_E - B - LO)
// and the thread's bits lie above TOP.  The stage's entry index
// lp = index >> (p+1) is then the thread's bits above TOP, th
// (nth = LOG_G - TOP of them), over the register bits [p+1, TOP)
// (nreg = TOP - 1 - p of them):
//     lp = th << nreg | rg.
// Lanes of a row differ in th, so in the table's own order they read entries
// 2^nreg apart, which strided 16-byte LDS reads serve with bank conflicts.
// The copy stores each stage's 2^j block as pos = rg << nth | th instead:
// for a given register the lanes read consecutive entries (and rows of one
// wave read the same ones).  Same entries, new positions.
template <int LOG_G, int LOG_E, bool FWD>
struct TwLayout {
    using Gm = Geom<LOG_G, LOG_E, FWD>;
    static constexpr int chunk_of(int p) {
        for (int c = 0; c < Gm::NC; ++c)
            if (p >= Gm::LO(c) && p < Gm::LO(c) + Gm::B(c)) return c;
        return 0;
    }
    static constexpr int top(int c) {
        const int spare = LOG_E - Gm::B(c) - Gm::LO(c);
        return Gm::LO(c) + Gm::B(c) + (spare > 0 ? spare : 0);
    }
    static constexpr int nth(int p) { return LOG_G - top(chunk_of(p)); }
    static constexpr int nreg(int p) { return LOG_G - 1 - p - nth(p); }
    // nth of every stage by j = LOG_G - 1 - p, four bits per j (LOG_G <= 16)
    static constexpr uint64_t nth_table() {
        uint64_t t = 0;
        for (int j = 0; j < LOG_G; ++j) t |= (uint64_t)nth(LOG_G - 1 - j) << (4 * j);
        return t;
    }
    // entry lp of stage j for block position pos (the inverse of the transposition)
    __device__ static __forceinline__ uint32_t lp_of(uint32_t j, uint32_t pos) {
        const uint32_t nt = (uint32_t)(nth_table() >> (4 * j)) & 15u;
        const uint32_t nr = j - nt;
        return ((pos & ((1u << nt) - 1u)) << nr) | (pos >> nt);
    }
};

template <int LOG_G, int LOG_E, bool FWD>
struct TwLdsT {
    using Lay = TwLayout<LOG_G, LOG_E, FWD>;
    const ulonglong2 *t;
    template <int P>
    __device__ __forceinline__ ulonglong2 get(uint32_t kt, uint32_t pr) const {
        constexpr int j = LOG_G - 1 - P;
        constexpr int nt = Lay::nth(P), nr = Lay::nreg(P);
        static_assert(nt + nr == j, "stage block size");
        // th = kt >> (P + 1 + nr) = kt >> TOP: the same shift for every stage of the chunk
        return t[tw_lds_index<LOG_G, FWD>(j, (pr << nt) | (kt >> (P + 1 + nr)))];
    }
};

// The chunk that holds a pass's top bits (a forward pass's chunk 0, an
// inverse pass's last chunk) has the thread's index bits below its register
// bits (Geom::kt(c, tid) < 2^LO(c)), so on each of its stages the entry
// (kt >> (p+1)) + pr is the compile-time pr alone.  Where the table range is
// also the same for the whole workgroup (a COL pass: one limb per
// workgroup; the stacked ROW pass: one row o and one plan per workgroup),
// that chunk's entries (chunk UC) are read through U with kt = 0: a uniform
// address, which the compiler serves with scalar loads instead of a
// per-lane LDS read each.  The other chunks read through T.
template <int UC, typename U, typename T>
struct TwTopChunkUniform {
    static constexpr bool kChunked = true;
    static constexpr int kUniformChunk = UC;
    U u;
    T t;
    template <int C, int P>
    __device__ __forceinline__ ulonglong2 getc(uint32_t kt, uint32_t pr) const {
        if constexpr (C == UC) return u.template get<P>(0u, pr);
        else return t.template get<P>(kt, pr);
    }
};

// The chunk that holds the top bits of a pass with LOG_G bits, LOG_E per chunk.
template <int LOG_G, int LOG_E, bool FWD>
constexpr int top_chunk() {
    return FWD ? 0 : Geom<LOG_G, LOG_E, FWD>::NC - 1;
}

template <typename TW, typename = void>
struct tw_chunked : std::false_type {};
template <typename TW>
struct tw_chunked<TW, std::void_t<decltype(TW::kChunked)>> : std::bool_constant<TW::kChunked> {};

// Staging of a pass's 2^LOG_G - 1 twiddle pairs into LDS by NT threads:
// fetch() issues the global loads into registers (K per thread), so the
// caller can issue its data loads before store() waits for them; the
// caller's barrier then publishes the copy.
template <int LOG_G, int NT>
struct TwStage {
    static constexpr uint32_t G = 1u << LOG_G;
    static constexpr int K = (int)((G - 1u + NT - 1) / NT);
    ulonglong2 v[K];

    // ROW pass, row o of degree n: entry i is (j, lp) of tw_lds_index, read
    // from the table at the stage base plus (o << j) + lp.
    // Entries past the end are clamped to the last one (a duplicate load,
    // not stored), keeping the loads branch-free so later loads need not
    // wait for them.
    // With R = n >> LOG_G rows, the table index of entry i folds into one
    // shift of a workgroup-uniform term:
    //     forward:  (R + o - 1) << j  +  i + 1                 j = floor(log2(i + 1))
    //     inverse:  (o + 2 - 2R) << j  +  i + n + 1 - G        j = floor(log2(G - 1 - i))
    // (the inverse's uniform term is negative for most rows; the sum is
    // formed modulo 2^32 and lands in [0, n)).
    // LOG_E >= 0 stages the copy in the transposed layout of TwLdsT for a
    // ROW pass with 2^LOG_E words per thread: slot i, position pos of its
    // stage's block, holds entry lp_of(j, pos), i.e. i + (lp - pos) in the
    // table order above.
    template <bool FWD, int LOG_E = -1>
    __device__ __forceinline__ void fetch_row(const ulonglong2 *tw, uint32_t n, uint32_t o, uint32_t t) {
        const uint32_t rows = n >> LOG_G;
        const uint32_t c = FWD ? rows + o - 1u : o + 2u - 2u * rows;  // uniform
        const uint32_t d = FWD ? 1u : n + 1u - G;                     // uniform
        static_for<0, K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const uint32_t i = min(t + (uint32_t)k * NT, G - 2u);
            const uint32_t j = 31u - __clz(FWD ? i + 1u : G - 1u - i);
            uint32_t e = i;
            if constexpr (LOG_E >= 0) {
                const uint32_t pos = FWD ? i + 1u - (1u << j) : i - (G - (2u << j));
                e = i - pos + TwLayout<LOG_G, LOG_E, FWD>::lp_of(j, pos);
            }
            v[k] = ld_g2(tw + ((c << j) + e + d));
        });
    }

    // COL pass: one contiguous range, entries [base, base + G - 1) with
    // base = 1 (forward) or n + 1 - G (inverse).
    template <bool FWD>
    __device__ __forceinline__ void fetch_col(const ulonglong2 *tw, uint32_t n, uint32_t t) {
        const uint32_t base = FWD ? 1u : n + 1u - G;
        static_for<0, K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            v[k] = ld_g2(tw + base + min(t + (uint32_t)k * NT, G - 2u));
        });
    }

    __device__ __forceinline__ void store(ulonglong2 *tw_lds, uint32_t t) const {
        static_for<0, K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const uint32_t i = t + (uint32_t)k * NT;
            if (i < G - 1u) tw_lds[i] = v[k];
        });
    }
};

// Stage kind of local bit p of a forward pass on the relaxed schedule
// (fwd_bfly_relaxed), s = LOG_G - 1 - p the stage's position in the pass:
//   COL pass (always the transform's first): inputs below 4q, so stages 0-5
//     skip the subtraction (words reach 16q), then one SUB8 in every four
//     (s = 6, 10, ...); it hands the next pass words below 16q;
//   ROW pass (second pass, or the only one): inputs below 16q, SUB8 at
//     s = 0, 4, 8, ..., and its bit 0, the transform's last stage, is LAST.
// Between two subtractions at most three SKIP stages follow a SUB8 (10q ->
// 16q), so every word stays below 16q.
template <int LOG_G, bool COL>
__device__ __forceinline__ constexpr FwdStage stage_kind(int p) {
    const int s = LOG_G - 1 - p;
    if (!COL && p == 0) return FwdStage::LAST;
    if (COL) return (s >= 6 && (s - 6) % 4 == 0) ? FwdStage::SUB8 : FwdStage::SKIP;
    return s % 4 == 0 ? FwdStage::SUB8 : FwdStage::SKIP;
}

// Orders one wave's LDS accesses for an exchange among its own lanes.  A
// wave's LDS instructions execute in issue order, so no hardware barrier is
// needed; the fences keep the compiler from moving a lane's LDS accesses
// across this point (it reasons per lane and would see no dependence).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Synchronization of a group's LDS exchange: the workgroup barrier, or
// wave_lds_sync when every group lies inside one wave and owns its LDS rows
// (WAVE).
template <bool WAVE>
__device__ __forceinline__ void group_sync() {
    if constexpr (WAVE) wave_lds_sync();
    else __syncthreads();
}

// A ROW group of T = 2^(LOG_G - LOG_E) threads lies inside one wave64 when T
// divides 64; its LDS words (Lds::idx2(g, ...)) are its own.
template <int LOG_G, int LOG_E>
constexpr bool row_group_in_wave() {
    return (64 % (1 << (LOG_G - LOG_E))) == 0;
}

// Moves registers through LDS: v[r] is written at word wi(r), then register
// r is read back from word ri(r) (wi, ri take the compile-time r).  SPLIT
// moves the words' 32-bit halves in two rounds through an array with the
// same word indices in 32-bit units: half the bytes, so more workgroups
// share a CU; the rounds after the first synchronize with group_sync<WAVE>.
// The synchronization between the writes and the first reads is
// group_sync<FIRST_WAVE> (a workgroup barrier where it must also publish
// other LDS data).  The caller orders the writes after earlier reads of the
// same words.
template <int E, bool SPLIT, bool WAVE, bool FIRST_WAVE, typename WI, typename RI>
__device__ __forceinline__ void lds_move(uint64_t (&v)[E], uint64_t *lds, WI wi, RI ri) {
    if constexpr (SPLIT) {
        uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
        uint32_t lo[E];
        static_for<0, E>([&](auto rc) { l32[wi(rc)] = (uint32_t)v[decltype(rc)::value]; });
        group_sync<FIRST_WAVE>();
        static_for<0, E>([&](auto rc) { lo[decltype(rc)::value] = l32[ri(rc)]; });
        group_sync<WAVE>();
        static_for<0, E>([&](auto rc) { l32[wi(rc)] = (uint32_t)(v[decltype(rc)::value] >> 32); });
        group_sync<WAVE>();
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            v[r] = (uint64_t)lo[r] | ((uint64_t)l32[ri(rc)] << 32);
        });
    } else {
        static_for<0, E>([&](auto rc) { lds[wi(rc)] = v[decltype(rc)::value]; });
        group_sync<FIRST_WAVE>();
        static_for<0, E>([&](auto rc) { v[decltype(rc)::value] = lds[ri(rc)]; });
    }
}

// The stacked ROW pass moves its words through LDS in 32-bit halves when its
// rows lie inside one wave (the extra rounds then synchronize one wave).
template <int LOG_G, int LOG_E>
constexpr bool stack_split() {
    return row_group_in_wave<LOG_G, LOG_E>();
}

// A COL pass moves its exchanges' words in 32-bit halves (lds_move; its
// groups span the workgroup, so every round takes a workgroup barrier),
// except on the relaxed forward schedule, where the extra barriers and LDS
// instructions cost more than the occupancy gives.
template <int LOG_G, int LOG_E, bool COL, bool RLX>
constexpr bool col_split() {
    return COL && !RLX && Geom<LOG_G, LOG_E, true>::NC > 1;
}

// Words (64-bit units) of the data tile in the LDS of ntt_pass_kernel: the
// tile, or its 32-bit halves rounded up to 16 bytes for a split COL pass.
template <int LOG_G, int LOG_E, int GPW, bool COL, bool RLX>
constexpr int pass_tile_words() {
    constexpr int w = Lds<LOG_G, LOG_E, GPW, COL>::WORDS;
    static_assert(w % 2 == 0, "16-byte aligned twiddle pairs");
    return col_split<LOG_G, LOG_E, COL, RLX>() ? (w + 3) / 4 * 2 : w;
}

// LDS of the stacked ROW pass in 64-bit words: the rows' words (their
// 32-bit halves when split, rounded up to 16 bytes), then the 2^LOG_G - 1
// twiddle pairs.
template <int LOG_G, int LOG_E, int GPW>
constexpr int stack_lds_words() {
    constexpr int w = Lds<LOG_G, LOG_E, GPW, false>::WORDS;
    static_assert(w % 2 == 0, "16-byte aligned twiddle pairs");
    return (stack_split<LOG_G, LOG_E>() ? (w + 3) / 4 * 2 : w) + 2 * ((1 << LOG_G) - 1);
}

// The butterfly chunks of one pass over registers v (chunk-0 layout on entry,
// last-chunk layout on exit), redistributing through LDS between chunks.
// Every thread of the workgroup must call it (it holds barriers).
// RLX: forward butterflies on the relaxed schedule (stage_kind).
// WSYNC: the exchanges synchronize each wave alone (ROW groups inside one
// wave, row_group_in_wave, whose LDS rows no other wave touches meanwhile).
// SPLIT: the exchanges move 32-bit halves (lds_move).
template <int LOG_G, int LOG_E, int GPW, bool FWD, bool COL, bool RLX = false, bool WSYNC = false,
          bool SPLIT = false, typename TW>
__device__ __forceinline__ void run_chunks(uint64_t (&v)[1 << LOG_E], uint64_t *lds, uint32_t g, uint32_t tid,
                                           const ModConsts &m, const TW &tws, bool final_inv, const NttJobDev &J) {
    using Gm = Geom<LOG_G, LOG_E, FWD>;
    using L = Lds<LOG_G, LOG_E, GPW, COL>;
    constexpr int E = 1 << LOG_E;
    constexpr int NC = Gm::NC;
    static_for<0, NC>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        constexpr int B = Gm::B(c);
        constexpr int LO = Gm::LO(c);
        const uint32_t ktid = Gm::kt(c, tid);

        if constexpr (c > 0) {
            // redistribute: registers in chunk c-1 layout -> LDS -> chunk c layout
            static_assert(!WSYNC || (!COL && row_group_in_wave<LOG_G, LOG_E>()), "wave-local exchange");
            const uint32_t kp = Gm::kt(c - 1, tid);
            group_sync<WSYNC>();
            lds_move<E, SPLIT, WSYNC, WSYNC>(
                v, lds, [&](auto rc) { return L::idx2(g, kp, Gm::kr(c - 1, decltype(rc)::value)); },
                [&](auto rc) { return L::idx2(g, ktid, Gm::kr(c, decltype(rc)::value)); });
        }

        // butterflies on local bits LO..LO+B-1 (forward: descending)
        static_for<0, B>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            constexpr int p = FWD ? LO + B - 1 - s : LO + s;
            constexpr int bit = 1 << (p - LO);
            const bool fin = !FWD && final_inv && (p == LOG_G - 1);
            uint64_t w[E], wq[E];
            static_for<0, E>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                if constexpr ((r & bit) == 0) {
                    constexpr int rp = Gm::rep(c, r, p);
                    constexpr uint32_t pr = Gm::kr(c, r) >> (p + 1);
                    if constexpr (rp == r) {
                        if (!fin) {
                            ulonglong2 tv;
                            if constexpr (tw_chunked<TW>::value) {
                                static_assert(Gm::LO(TW::kUniformChunk) + Gm::B(TW::kUniformChunk) == LOG_G,
                                              "only the top chunk's entries are uniform");
                                tv = tws.template getc<c, p>(ktid, pr);
                            } else {
                                tv = tws.template get<p>(ktid, pr);
                            }
                            w[r] = tv.x;
                            wq[r] = tv.y;
                        }
                    }
                    if constexpr (FWD && RLX) {
                        constexpr FwdStage S = stage_kind<LOG_G, COL>(p);
                        fwd_bfly_relaxed<S>(v[r], v[r | bit], w[rp], wq[rp], m);
                    } else if constexpr (FWD) {
                        fwd_bfly(v[r], v[r | bit], w[rp], wq[rp], m);
                    } else {
                        if (fin)
                            inv_final_bfly(v[r], v[r | bit], J.ninv, J.ninv_q, J.sroot, J.sroot_q, m);
                        else
                            inv_bfly(v[r], v[r | bit], w[rp], wq[rp], m);
                    }
                }
            });
        });
    });
}

template <int LOG_G, int LOG_E, int GPW, bool FWD, bool COL, bool MULTI, bool RLX = false>
__global__ void __launch_bounds__(GPW << (LOG_G - LOG_E))
ntt_pass_kernel(PassArgs a) {
    using Gm = Geom<LOG_G, LOG_E, FWD>;
    using L = Lds<LOG_G, LOG_E, GPW, COL>;
    constexpr int G = 1 << LOG_G;
    constexpr int E = 1 << LOG_E;
    constexpr int T = G / E;
    constexpr int NT = GPW * T;
    constexpr int NC = Gm::NC;
    extern __shared__ uint64_t lds[];

    // XCD-aware tile order: consecutive workgroup ids are spread round-robin
    // over the 8 XCDs, so give each XCD a contiguous slice of the tile space
    // (limbs sharing a twiddle table then share that XCD's L2).
    uint32_t tile = blockIdx.x;
    {
        const uint32_t nb = gridDim.x;
        if ((nb & 7u) == 0) tile = (tile & 7u) * (nb >> 3) + (tile >> 3);
    }

    const uint32_t t = threadIdx.x;
    uint32_t g, tid;
    if constexpr (COL) {
        g = t % GPW;
        tid = t / GPW;
    } else {
        tid = t % T;
        g = t / T;
    }

    // Group -> (limb, group-in-limb); groups_per_limb = N >> LOG_G = 2^lgpl.
    const uint32_t lgpl = a.logn - LOG_G;
    const uint32_t gid = tile * GPW + g;
    uint32_t limb, o;
    if constexpr (MULTI) {
        limb = gid >> lgpl;
        o = gid & ((1u << lgpl) - 1u);
    } else {
        limb = (tile * GPW) >> lgpl;  // uniform over the workgroup
        o = gid - (limb << lgpl);
    }
    const bool valid = limb < a.num_limbs;
    const uint32_t jl = valid ? limb : 0;
    const NttJobDev J = load_job(a, jl);
    const ModConsts m = mod_consts(J.q);
    const uint64_t *src = a.first_pass ? J.src : J.dst;
    uint64_t *dst = J.dst;
    const ulonglong2 *tw = J.tw;
    const uint32_t logn = a.logn;
    const uint32_t n = 1u << logn;
    const uint32_t blo = COL ? logn - LOG_G : 0u;

    // Word offset of local index k of this thread's group.
    auto gaddr = [&](uint32_t k) -> uint32_t {
        if constexpr (COL) return (k << blo) | o;
        else return (o << LOG_G) | k;
    };
    // ROW tiles of one limb are GPW * G contiguous words.
    const uint32_t tile_o0 = MULTI ? 0u : tile * GPW - (limb << lgpl);

    uint64_t v[E];

    // ---- COL passes: stage the pass's twiddles in LDS (one contiguous range
    // per pass, shared by every column of the limb); the barrier below the
    // data loads publishes them.
    ulonglong2 *tw_lds = reinterpret_cast<ulonglong2 *>(lds + pass_tile_words<LOG_G, LOG_E, GPW, COL, RLX>());
    TwStage<LOG_G, NT> tstage;
    if constexpr (COL) tstage.template fetch_col<FWD>(tw, n, t);

    // ---- load: registers in chunk-0 layout
    constexpr bool LOAD_VIA_LDS = !COL && !FWD;  // inverse ROW: chunk 0 is thread-contiguous
    if constexpr (LOAD_VIA_LDS) {
        if constexpr (!MULTI) {
            // contiguous tile -> LDS, 16-byte vectors when aligned
            const uint64_t *ts = src + ((uint64_t)tile_o0 << LOG_G);
            if (a.vec16) {
#pragma unroll
                for (int i = 0; i < E / 2; ++i) {
                    const uint32_t e = 2u * (t + (uint32_t)i * NT);
                    const ulonglong2 x2 = ld_nt2(ts + e);
                    const uint32_t ge = e >> LOG_G, ke = e & (G - 1);
                    lds[L::idx(ge, ke)] = x2.x;
                    lds[L::idx(ge, ke + 1)] = x2.y;
                }
            } else {
#pragma unroll
                for (int i = 0; i < E; ++i) {
                    const uint32_t e = t + (uint32_t)i * NT;
                    lds[L::idx(e >> LOG_G, e & (G - 1))] = ld_g(ts + e);
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const uint32_t e = t + (uint32_t)i * NT;
                const uint32_t ge = e >> LOG_G, ke = e & (G - 1);
                const uint32_t le = tile * GPW + ge;  // one group per limb
                if (le < a.num_limbs) lds[L::idx(ge, ke)] = ld_g(job_src(a, le) + ke);
            }
        }
        __syncthreads();
        const uint32_t k0 = Gm::kt(0, tid);
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            v[r] = lds[L::idx2(g, k0, Gm::kr(0, r))];
        });
    } else {
        const uint32_t k0 = Gm::kt(0, tid);
        if (valid) {
            if constexpr (COL) {
                // word (k0 + kr) << blo | o: the register part kr << blo is
                // uniform (scalar base), the thread part one 32-bit offset
                const uint32_t toff = ((k0 << blo) | o) << 3;
                static_for<0, E>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    v[r] = ld_gb(src + ((uint64_t)Gm::kr(0, r) << blo), toff);
                });
            } else {
                static_for<0, E>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    v[r] = ld_g(src + gaddr(k0 + Gm::kr(0, r)));
                });
            }
        } else {
            static_for<0, E>([&](auto rc) { v[decltype(rc)::value] = 0; });
        }
    }

    // ---- chunks
    if constexpr (COL) {
        tstage.store(tw_lds, t);
        __syncthreads();
        // one limb per workgroup: the top chunk's entries are uniform
        const TwTopChunkUniform<top_chunk<LOG_G, LOG_E, FWD>(), TwGlobal<LOG_G, FWD, true>, TwLds<LOG_G, FWD>> tws{
            {tw, n, blo, o}, {tw_lds}};
        run_chunks<LOG_G, LOG_E, GPW, FWD, COL, RLX, false, col_split<LOG_G, LOG_E, COL, RLX>()>(
            v, lds, g, tid, m, tws, a.final_inv != 0, J);
    } else {
        // a ROW group's exchanges touch its own LDS words only (the staged
        // loads above and the staged stores below are published by
        // workgroup barriers): a group inside one wave synchronizes alone
        const TwGlobal<LOG_G, FWD, COL> tws{tw, n, blo, o};
        run_chunks<LOG_G, LOG_E, GPW, FWD, COL, RLX, row_group_in_wave<LOG_G, LOG_E>()>(v, lds, g, tid, m, tws,
                                                                                      a.final_inv != 0, J);
    }

    // ---- normalisation folded into the last pass (ntt.cpp:79-118)
    if (a.normalize) {
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (FWD) v[r] = csub_n(csub_n(v[r], m.q2, m.nq2), m.q, m.nq);
            else v[r] = csub_n(v[r], m.q, m.nq);
        });
    }

    // ---- store
    constexpr bool STORE_VIA_LDS = !COL && FWD;  // forward ROW: last chunk is thread-contiguous
    if constexpr (STORE_VIA_LDS) {
        const uint32_t kl = Gm::kt(NC - 1, tid);
        __syncthreads();
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            lds[L::idx2(g, kl, Gm::kr(NC - 1, r))] = v[r];
        });
        __syncthreads();
        if constexpr (!MULTI) {
            uint64_t *td = dst + ((uint64_t)tile_o0 << LOG_G);
            if (a.vec16) {
#pragma unroll
                for (int i = 0; i < E / 2; ++i) {
                    const uint32_t e = 2u * (t + (uint32_t)i * NT);
                    const uint32_t ge = e >> LOG_G, ke = e & (G - 1);
                    ulonglong2 x2;
                    x2.x = lds[L::idx(ge, ke)];
                    x2.y = lds[L::idx(ge, ke + 1)];
                    st_g2(td + e, x2);
                }
            } else {
#pragma unroll
                for (int i = 0; i < E; ++i) {
                    const uint32_t e = t + (uint32_t)i * NT;
                    st_g(td + e, lds[L::idx(e >> LOG_G, e & (G - 1))]);
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const uint32_t e = t + (uint32_t)i * NT;
                const uint32_t ge = e >> LOG_G, ke = e & (G - 1);
                const uint32_t le = tile * GPW + ge;
                if (le < a.num_limbs) st_g(job_dst(a, le) + ke, lds[L::idx(ge, ke)]);
            }
        }
    } else {
        if (valid) {
            const uint32_t kl = Gm::kt(NC - 1, tid);
            if constexpr (COL) {
                const uint32_t toff = ((kl << blo) | o) << 3;
                static_for<0, E>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    st_gb(dst + ((uint64_t)Gm::kr(NC - 1, r) << blo), toff, v[r]);
                });
            } else {
                static_for<0, E>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    st_g(dst + gaddr(kl + Gm::kr(NC - 1, r)), v[r]);
                });
            }
        }
    }
}

// ROW pass over stacks of jobs that share one plan (same modulus and
// twiddle table).  Workgroup (stack s, row o) transforms row o of the GPW
// jobs jobs[s*GPW .. s*GPW+GPW-1]; an empty slot has dst == nullptr, slot 0
// is always a job.  Every row of the workgroup uses the same 2^LOG_G - 1
// twiddle pairs, which are staged in LDS once (for the per-limb ROW pass
// each row needs its own set, twice the row's own bytes, from L2).  Bits,
// butterflies and twiddles are those of ntt_pass_kernel's ROW pass.
template <int LOG_G, int LOG_E, int GPW, bool FWD, bool RLX = false>
__global__ void __launch_bounds__(GPW << (LOG_G - LOG_E)) __attribute__((amdgpu_waves_per_eu(RLX ? 6 : 1)))
ntt_row_stack_kernel(PassArgs a) {
    using Gm = Geom<LOG_G, LOG_E, FWD>;
    using L = Lds<LOG_G, LOG_E, GPW, false>;
    constexpr int G = 1 << LOG_G;
    constexpr int E = 1 << LOG_E;
    constexpr int T = G / E;
    constexpr int NT = GPW * T;
    constexpr int NC = Gm::NC;
    // a row's LDS words are its own: when a row lies inside one wave, its
    // exchanges and its load or store staging synchronize that wave alone,
    // and the words move in 32-bit halves through half the LDS
    constexpr bool WS = row_group_in_wave<LOG_G, LOG_E>();
    constexpr bool SPLIT = stack_split<LOG_G, LOG_E>();
    extern __shared__ uint64_t lds[];

    uint32_t tile = blockIdx.x;
    {
        const uint32_t nb = gridDim.x;
        if ((nb & 7u) == 0) tile = (tile & 7u) * (nb >> 3) + (tile >> 3);
    }
    const uint32_t t = threadIdx.x;
    const uint32_t tid = t % T, g = t / T;
    const uint32_t lgpl = a.logn - LOG_G;        // groups_per_limb = 2^lgpl
    const uint32_t st = tile >> lgpl;            // stack (num_limbs = number of stacks)
    const uint32_t o = tile & ((1u << lgpl) - 1u);  // row
    const NttJobDev J0 = a.jobs[st * GPW];              // the stack's plan
    const NttJobDev *Jg = a.jobs + st * GPW + g;
    uint64_t *dst = Jg->dst;
    const uint64_t *src = a.first_pass ? Jg->src : dst;
    const bool valid = dst != nullptr;
    const ModConsts m = mod_consts(J0.q);
    const uint32_t n = 1u << a.logn;

    ulonglong2 *tw_lds = reinterpret_cast<ulonglong2 *>(lds + (stack_lds_words<LOG_G, LOG_E, GPW>() - 2 * (G - 1)));
    TwStage<LOG_G, NT> tstage;
    tstage.template fetch_row<FWD, LOG_E>(J0.tw, n, o, t);

    // an empty slot reads slot 0's row (always a job) and stores nothing:
    // no separate path for its registers
    const uint64_t *lsrc = valid ? src : (a.first_pass ? J0.src : J0.dst);
    const uint64_t row = (uint64_t)o << LOG_G;
    // word (2 tid + (r & 1), (r >> 1) 2T) of the row: register r of the
    // thread-contiguous 16-byte units in which the inverse loads a row and the
    // forward stores one (the register part 2T (r >> 1) is disjoint from the
    // thread part 2 tid + 1)
    auto unit_word = [&](auto rc) {
        constexpr int r = decltype(rc)::value;
        return L::idx2(g, 2u * tid + (uint32_t)(r & 1), 2u * (uint32_t)(r >> 1) * T);
    };
    uint64_t v[E];
    if constexpr (!FWD) {
        // inverse: chunk 0 is thread-contiguous; load the row in 16-byte
        // units and redistribute them through LDS
        if (a.vec16) {
#pragma unroll
            for (int i = 0; i < E / 2; ++i) {
                const ulonglong2 x2 = ld_nt2(lsrc + row + 2u * tid + 2u * (uint32_t)i * T);
                v[2 * i] = x2.x;
                v[2 * i + 1] = x2.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i)
                v[i] = ld_g(lsrc + row + 2u * tid + (uint32_t)(i & 1) + 2u * (uint32_t)(i >> 1) * T);
        }
        tstage.store(tw_lds, t);
        const uint32_t k0 = Gm::kt(0, tid);
        // the first barrier is the workgroup's: it also publishes the twiddles
        lds_move<E, SPLIT, WS, false>(v, lds, unit_word,
                                      [&](auto rc) { return L::idx2(g, k0, Gm::kr(0, decltype(rc)::value)); });
    } else {
        const uint32_t k0 = Gm::kt(0, tid);
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            v[r] = ld_g(lsrc + row + k0 + Gm::kr(0, r));
        });
        tstage.store(tw_lds, t);
        __syncthreads();  // publishes the twiddles
    }

    // one row o and one plan per workgroup: the top chunk's entries are uniform
    const TwTopChunkUniform<top_chunk<LOG_G, LOG_E, FWD>(), TwGlobal<LOG_G, FWD, false>, TwLdsT<LOG_G, LOG_E, FWD>> tws{
        {J0.tw, n, 0u, o}, {tw_lds}};
    run_chunks<LOG_G, LOG_E, GPW, FWD, false, RLX, WS, SPLIT>(v, lds, g, tid, m, tws, false, J0);

    if (a.normalize) {
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (FWD) v[r] = csub_n(csub_n(v[r], m.q2, m.nq2), m.q, m.nq);
            else v[r] = csub_n(v[r], m.q, m.nq);
        });
    }

    const uint32_t kl = Gm::kt(NC - 1, tid);
    if constexpr (FWD) {
        // forward: the last chunk is thread-contiguous; store the row in
        // 16-byte units redistributed through LDS
        group_sync<WS>();
        lds_move<E, SPLIT, WS, WS>(
            v, lds, [&](auto rc) { return L::idx2(g, kl, Gm::kr(NC - 1, decltype(rc)::value)); }, unit_word);
        if (valid) {
            if (a.vec16) {
#pragma unroll
                for (int i = 0; i < E / 2; ++i) {
                    ulonglong2 x2;
                    x2.x = v[2 * i];
                    x2.y = v[2 * i + 1];
                    st_g2(dst + row + 2u * tid + 2u * (uint32_t)i * T, x2);
                }
            } else {
#pragma unroll
                for (int i = 0; i < E; ++i)
                    st_g(dst + row + 2u * tid + (uint32_t)(i & 1) + 2u * (uint32_t)(i >> 1) * T, v[i]);
            }
        }
    } else if (valid) {
        static_for<0, E>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            st_g(dst + row + kl + Gm::kr(NC - 1, r), v[r]);
        });
    }
}

}  // namespace axhel_dev
