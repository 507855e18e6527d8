// RNS context: N and L primes with resident plans, and calls over whole RNS
// polynomials stored [count][L][N] in device memory.  The reference leaves
// the limb loop to its caller (one ComputeForward / Eltwise*Mod call per
// limb, integration/seal/SEAL-4.4.5-AXHEL.patch:185-226); here one call
// covers count x L limbs.
//
// Each call builds its job table on the device (a fill kernel), so no table
// crosses PCIe, into a table buffer the context keeps per stream: calls on
// one stream run in order, so the buffer is free again for the next call on
// that stream, and calls on different streams never share one.  A call that
// repeats the previous call on its stream (NTT: same buffers, count,
// direction and stack layout; element-wise: same buffers, count and per-limb
// scalars) finds its table in place and skips the fill.  The context's
// mutex is held while a call enqueues its work, so two threads issuing on
// one stream cannot interleave a fill with another call's passes.  The
// transforms then run through the batched kernels, the ROW pass stacked over
// polynomials when the stacks fill.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "eltwise_host.hpp"
#include "ntt_host.hpp"

using axhel_dev::Barrett;
using axhel_dev::NttJobDev;
using namespace axhel_host;

namespace {

constexpr uint32_t kMaxLimbs = AXHEL_HIP_RNS_MAX_LIMBS;  // per-limb scalars travel in the fill kernel's arguments

struct RnsLimbDev {
    const ulonglong2 *twf, *twi;
    uint64_t q, ninv, ninv_q, sroot, sroot_q;
    Barrett bc;
};

}  // namespace

struct axhel_hip_rns {
    uint64_t n = 0;
    uint32_t logn = 0;
    uint32_t limbs = 0;
    std::vector<axhel_hip_plan *> plans;
    RnsLimbDev *d_info = nullptr;
    bool relax_ok = true;  // every modulus below 2^60
    // job-table buffer per stream, and the NTT call whose table it holds
    struct StreamTable {
        hipStream_t stream;
        void *buf;
        size_t bytes;
        bool ntt_valid;
        const void *result, *operand;
        uint64_t count;
        int inverse;
        uint32_t slots;  // stack slots per limb (0: no stack table)
        // or the element-wise call whose table it holds (the op and factor
        // are launch parameters, not table content)
        bool elt_valid;
        const void *e_r, *e_a, *e_b;
        uint64_t e_count;
        uint64_t e_s[AXHEL_HIP_RNS_MAX_LIMBS];
    };
    std::mutex mu;
    std::vector<StreamTable> tables;
};

namespace {

typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        e >>= 1;
        if (e) b = mulmod(b, b, q);
    }
    return r;
}

// The reference tests' root rule (parameters.hpp:110-122): the first
// a = 2, 3, ... whose a^((q-1)/2N) has order exactly 2N.
bool find_root(uint64_t n, uint64_t q, uint64_t *root) {
    if ((q - 1) % (2 * n) != 0) return false;
    const uint64_t e = (q - 1) / (2 * n);
    for (uint64_t a = 2; a < 100000 && a < q; ++a) {
        const uint64_t psi = powmod(a, e, q);
        if (psi != 1 && powmod(psi, n, q) == q - 1 && powmod(psi, 2 * n, q) == 1) {
            *root = psi;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ NttJobDev job_for(const RnsLimbDev &L, const uint64_t *src, uint64_t *dst, bool inverse) {
    NttJobDev j;
    j.src = src;
    j.dst = dst;
    j.tw = inverse ? L.twi : L.twf;
    j.q = L.q;
    j.ninv = L.ninv;
    j.ninv_q = L.ninv_q;
    j.sroot = L.sroot;
    j.sroot_q = L.sroot_q;
    return j;
}

// Job tables of one transform call.  jobs: limb-major, entry l*count + p is
// limb l of polynomial p (so jobs sharing a plan are adjacent).  stacks (if
// non-null): per limb, ceil(count/gpw) stacks of gpw slots over polynomials
// s*gpw + g, empty slots (dst null) past count.
__global__ void rns_fill_ntt(NttJobDev *jobs, NttJobDev *stacks, const RnsLimbDev *info, uint64_t *res,
                             const uint64_t *op, uint32_t count, uint32_t limbs, uint64_t n, int inverse,
                             uint32_t slots_per_limb) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nj = (uint64_t)count * limbs;
    if (t < nj) {
        const uint32_t l = (uint32_t)(t / count), p = (uint32_t)(t % count);
        const uint64_t off = ((uint64_t)p * limbs + l) * n;
        jobs[t] = job_for(info[l], op + off, res + off, inverse != 0);
    }
    if (stacks && t < (uint64_t)slots_per_limb * limbs) {
        const uint32_t l = (uint32_t)(t / slots_per_limb), p = (uint32_t)(t % slots_per_limb);
        if (p < count) {
            const uint64_t off = ((uint64_t)p * limbs + l) * n;
            stacks[t] = job_for(info[l], op + off, res + off, inverse != 0);
        } else {
            stacks[t] = job_for(info[l], nullptr, nullptr, inverse != 0);
        }
    }
}

struct EltFillArgs {
    uint64_t s[kMaxLimbs];  // per-limb scalar (already reduced as the op requires)
};

__global__ void rns_fill_elt(EltLimbDev *out, const RnsLimbDev *info, uint64_t *r, const uint64_t *a,
                             const uint64_t *b, uint32_t count, uint32_t limbs, uint64_t n, EltFillArgs sc) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)count * limbs) return;
    const uint32_t l = (uint32_t)(t % limbs);
    EltLimbDev d;
    d.r = r + t * n;  // t = p * limbs + l: polynomial-major, as stored
    d.a = a + t * n;
    d.b = b ? b + t * n : nullptr;
    d.s = sc.s[l];
    d.bc = info[l].bc;
    out[t] = d;
}

bool device_ptr(const void *p) { return p && is_device_pointer(p); }

// The context's table buffer for stream st, at least `bytes` long (caller
// holds ctx->mu).  Growing waits for the stream's queued work, which may
// still read the old buffer, before freeing it; buffers only grow.
hipError_t stream_table(axhel_hip_rns *ctx, hipStream_t st, size_t bytes, axhel_hip_rns::StreamTable **out) {
    axhel_hip_rns::StreamTable *t = nullptr;
    for (auto &e : ctx->tables)
        if (e.stream == st) t = &e;
    if (!t) {
        ctx->tables.emplace_back();
        t = &ctx->tables.back();
        *t = axhel_hip_rns::StreamTable{};
        t->stream = st;
    }
    if (t->bytes < bytes) {
        if (t->buf) {
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            (void)hipFree(t->buf);
            t->buf = nullptr;
            t->bytes = 0;
        }
        t->ntt_valid = false;
        t->elt_valid = false;
        const size_t want = std::max(bytes, size_t(64) << 10);
        hipError_t e = hipMalloc(&t->buf, want);
        if (e != hipSuccess) return e;
        t->bytes = want;
    }
    *out = t;
    return hipSuccess;
}

// Calls on a capturing stream are refused: a captured call would depend on
// the context's per-stream table buffer, which later calls overwrite (and a
// call that skips the fill would record no table at all).  Batches are the
// capturable form of the same work.
axhel_hip_status refuse_capture(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const hipError_t e = hipStreamIsCapturing(st, &cs);
    if (e != hipSuccess) {
        set_last_error("rns: stream state", e);
        return AXHEL_HIP_HIP_ERROR;
    }
    if (cs != hipStreamCaptureStatusNone) {
        set_last_error_text("rns: calls cannot be captured into a graph; capture a batch (axhel_hip_ntt_batch_*)");
        return AXHEL_HIP_INVALID_ARGUMENT;
    }
    return AXHEL_HIP_OK;
}

// ---- no-allocation form of the NTT call
//
// The job and stack tables of axhel_hip_rns_ntt_noalloc live in this
// file-scope device array instead of a per-stream buffer the context
// allocates.  Each GPU holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: the
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is passed to kernels.  Every
// call on a GPU writes the same array, so calls that overlap on one GPU
// (concurrent streams, several threads) race.
constexpr size_t kScratchEntries = AXHEL_HIP_RNS_NOALLOC_MAX_ENTRIES;
static __device__ __align__(256) NttJobDev scratch[kScratchEntries];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
NttJobDev *g_scratch_dev[kMaxDevices];

// The scratch array's address on the current device (set-up on first use).
hipError_t scratch_for_device(NttJobDev **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[dev] = static_cast<NttJobDev *>(p);
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

// Stack layout of a transform call over count x L limbs: slots per limb of
// the stacked ROW pass (0: no stack table), the same rule as the batches
// (padding each limb's polynomials to whole stacks leaves at most one slot
// in eight empty).
uint32_t stack_slots_per_limb(const NttConfig &cfg, uint32_t cnt, uint32_t L) {
    const int gpw = stack_gpw_of(cfg);
    if (!gpw) return 0;
    const uint64_t nj = (uint64_t)cnt * L;
    const uint64_t spl = (cnt + (uint64_t)gpw - 1) / gpw * gpw;
    return (spl * L - nj) * 8 <= spl * L ? (uint32_t)spl : 0;
}

// Fill, then the passes, of one transform call whose tables go to `jobs`
// (nj entries, then the stack table when slots_per_limb != 0).
hipError_t enqueue_rns_ntt(const axhel_hip_rns *ctx, const NttConfig &cfg, NttJobDev *jobs, bool fill,
                           uint64_t *result, const uint64_t *operand, uint32_t cnt, int inverse, bool normalize,
                           uint32_t slots_per_limb, hipStream_t st) {
    const uint32_t L = ctx->limbs;
    const uint64_t nj = (uint64_t)cnt * L;
    NttJobDev *stacks = slots_per_limb ? jobs + nj : nullptr;
    if (fill) {
        const uint64_t n_fill = std::max<uint64_t>(nj, (uint64_t)slots_per_limb * L);
        hipLaunchKernelGGL(rns_fill_ntt, dim3((unsigned)((n_fill + 255) / 256)), dim3(256), 0, st, jobs, stacks,
                           ctx->d_info, result, operand, cnt, L, ctx->n, inverse, slots_per_limb);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    static const NttJobDev none{};
    StackTable tab;
    if (stacks) {
        const int gpw = stack_gpw_of(cfg);
        tab.jobs = stacks;
        tab.count = slots_per_limb * L / gpw;
        tab.gpw = gpw;
    }
    const bool vec16 = aligned16(result) && aligned16(operand);
    return launch_ntt(cfg, ctx->logn, inverse != 0, normalize, jobs, none, (uint32_t)nj, st, -1, vec16,
                      stacks ? &tab : nullptr, ctx->relax_ok);
}

}  // namespace

extern "C" {

axhel_hip_status axhel_hip_rns_create(uint64_t degree, const uint64_t *moduli, const uint64_t *roots, uint32_t limbs,
                                      axhel_hip_rns **ctx) {
    if (!ctx || !moduli || limbs == 0 || limbs > kMaxLimbs) return AXHEL_HIP_INVALID_ARGUMENT;
    *ctx = nullptr;
    // the degree rule of plan creation, checked before the root search divides by 2N
    if (degree < 2 || (degree & (degree - 1)) != 0 || degree > (1ull << kMaxLogN)) return AXHEL_HIP_INVALID_ARGUMENT;
    axhel_hip_rns *c = new (std::nothrow) axhel_hip_rns;
    if (!c) return AXHEL_HIP_NO_MEMORY;
    c->n = degree;
    c->limbs = limbs;
    std::vector<RnsLimbDev> info(limbs);
    axhel_hip_status st = AXHEL_HIP_OK;
    for (uint32_t l = 0; l < limbs && st == AXHEL_HIP_OK; ++l) {
        uint64_t psi = roots ? roots[l] : 0;
        if (!roots && !find_root(degree, moduli[l], &psi)) {
            set_last_error("rns create: no primitive 2N-th root for a modulus", hipErrorInvalidValue);
            st = AXHEL_HIP_INVALID_ARGUMENT;
            break;
        }
        axhel_hip_plan *p = nullptr;
        st = axhel_hip_ntt_plan_create(degree, moduli[l], psi, &p);
        if (st != AXHEL_HIP_OK) break;
        c->plans.push_back(p);
        c->logn = p->logn;
        c->relax_ok = c->relax_ok && relax_ok_modulus(p->q);
        info[l] = RnsLimbDev{p->d_fwd, p->d_inv, p->q, p->ninv, p->ninv_q, p->sroot, p->sroot_q, barrett_for(p->q)};
    }
    if (st == AXHEL_HIP_OK) {
        hipError_t e = hipMalloc(&c->d_info, limbs * sizeof(RnsLimbDev));
        if (e == hipSuccess) e = hipMemcpy(c->d_info, info.data(), limbs * sizeof(RnsLimbDev), hipMemcpyHostToDevice);
        // set-up of the no-allocation call on this device: the scratch
        // array's address and the degree's kernel attributes
        NttJobDev *scr = nullptr;
        if (e == hipSuccess) e = scratch_for_device(&scr);
        if (e == hipSuccess) e = prepare_config(config_for_single(c->logn));
        if (e == hipSuccess) e = prepare_config(config_for_small(c->logn));
        if (e == hipSuccess) e = prepare_config(config_for(c->logn));
        if (e != hipSuccess) {
            set_last_error("rns create", e);
            st = e == hipErrorOutOfMemory ? AXHEL_HIP_NO_MEMORY : AXHEL_HIP_HIP_ERROR;
        }
    }
    if (st != AXHEL_HIP_OK) {
        axhel_hip_rns_destroy(c);
        return st;
    }
    *ctx = c;
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_rns_destroy(axhel_hip_rns *ctx) {
    if (!ctx) return AXHEL_HIP_OK;
    for (auto &t : ctx->tables) (void)hipFree(t.buf);
    for (auto *p : ctx->plans) axhel_hip_ntt_plan_destroy(p);
    (void)hipFree(ctx->d_info);
    delete ctx;
    return AXHEL_HIP_OK;
}

const axhel_hip_plan *axhel_hip_rns_plan(const axhel_hip_rns *ctx, uint32_t limb) {
    return ctx && limb < ctx->limbs ? ctx->plans[limb] : nullptr;
}

axhel_hip_status axhel_hip_rns_ntt(axhel_hip_rns *ctx, uint64_t *result, const uint64_t *operand, uint64_t count,
                                   int inverse, uint64_t output_mod_factor, void *stream) {
    if (!ctx || (inverse != 0 && inverse != 1)) return AXHEL_HIP_INVALID_ARGUMENT;
    if (count == 0) return AXHEL_HIP_OK;
    if (!device_ptr(result) || !device_ptr(operand) || count > 0x7FFFFFFFull / ctx->limbs)
        return AXHEL_HIP_INVALID_ARGUMENT;
    // lazy only for output factor 4 (forward) / 2 (inverse), as ntt.cpp:173-204
    const bool normalize = inverse ? output_mod_factor != 2 : output_mod_factor != 4;
    const hipStream_t st = as_stream(stream);
    if (const axhel_hip_status cs = refuse_capture(st)) return cs;
    const uint32_t cnt = (uint32_t)count, L = ctx->limbs;
    const uint64_t nj = (uint64_t)cnt * L;
    const NttConfig &cfg = config_for_launch(ctx->logn, nj);
    AXHEL_HIP_TRY(prepare_config(cfg));
    const uint32_t slots_per_limb = stack_slots_per_limb(cfg, cnt, L);
    const uint64_t entries = nj + (uint64_t)slots_per_limb * L;
    std::lock_guard<std::mutex> lk(ctx->mu);
    axhel_hip_rns::StreamTable *tab_buf = nullptr;
    AXHEL_HIP_TRY(stream_table(ctx, st, entries * sizeof(NttJobDev), &tab_buf));
    // the table depends on the buffers, the count, the direction and the
    // stack layout (slots per limb, from the configuration's group count): a
    // repeated call finds it in place
    const bool same = tab_buf->ntt_valid && tab_buf->result == result && tab_buf->operand == operand &&
                      tab_buf->count == count && tab_buf->inverse == inverse && tab_buf->slots == slots_per_limb;
    if (!same) {
        tab_buf->ntt_valid = false;
        tab_buf->elt_valid = false;  // the NTT table overwrites the buffer
    }
    const hipError_t e = enqueue_rns_ntt(ctx, cfg, (NttJobDev *)tab_buf->buf, !same, result, operand, cnt, inverse,
                                         normalize, slots_per_limb, st);
    if (e == hipSuccess && !same) {
        tab_buf->ntt_valid = true;
        tab_buf->result = result;
        tab_buf->operand = operand;
        tab_buf->count = count;
        tab_buf->inverse = inverse;
        tab_buf->slots = slots_per_limb;
    }
    AXHEL_HIP_TRY(e);
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_rns_ntt_noalloc(axhel_hip_rns *ctx, uint64_t *result, const uint64_t *operand,
                                           uint64_t count, int inverse, uint64_t output_mod_factor, void *stream) {
    if (!ctx || (inverse != 0 && inverse != 1)) return AXHEL_HIP_INVALID_ARGUMENT;
    if (count == 0) return AXHEL_HIP_OK;
    if (!device_ptr(result) || !device_ptr(operand) || count > 0x7FFFFFFFull / ctx->limbs)
        return AXHEL_HIP_INVALID_ARGUMENT;
    const bool normalize = inverse ? output_mod_factor != 2 : output_mod_factor != 4;
    const hipStream_t st = as_stream(stream);
    if (const axhel_hip_status cs = refuse_capture(st)) return cs;
    const uint32_t cnt = (uint32_t)count, L = ctx->limbs;
    const uint64_t nj = (uint64_t)cnt * L;
    const NttConfig &cfg = config_for_launch(ctx->logn, nj);
    const uint32_t slots_per_limb = stack_slots_per_limb(cfg, cnt, L);
    const uint64_t entries = nj + (uint64_t)slots_per_limb * L;
    if (entries > kScratchEntries) {
        set_last_error_text("rns_ntt_noalloc: count x limbs needs " + std::to_string(entries) +
                            " job descriptors, above the scratch maximum of " + std::to_string(kScratchEntries));
        return AXHEL_HIP_INVALID_ARGUMENT;
    }
    NttJobDev *scr = nullptr;
    AXHEL_HIP_TRY(scratch_for_device(&scr));
    // kernel attributes were raised at set-up for the selected configurations;
    // this only finds them recorded
    AXHEL_HIP_TRY(prepare_config(cfg));
    std::lock_guard<std::mutex> lk(ctx->mu);
    AXHEL_HIP_TRY(enqueue_rns_ntt(ctx, cfg, scr, true, result, operand, cnt, inverse, normalize, slots_per_limb, st));
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_rns_eltwise(axhel_hip_rns *ctx, int op, uint64_t *result, const uint64_t *operand1,
                                       const uint64_t *operand2, const uint64_t *scalars, uint64_t count,
                                       uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream) {
    if (!ctx) return AXHEL_HIP_INVALID_ARGUMENT;
    if (count == 0) return AXHEL_HIP_OK;
    if (!device_ptr(result) || !device_ptr(operand1) || (operand2 && !device_ptr(operand2)) ||
        count > 0x7FFFFFFFull / ctx->limbs)
        return AXHEL_HIP_INVALID_ARGUMENT;
    const uint32_t L = ctx->limbs;
    const hipStream_t st = as_stream(stream);
    if (const axhel_hip_status cs = refuse_capture(st)) return cs;
    EltFillArgs sc{};
    int eop = 0, f = 1;
    const uint64_t *b = nullptr;
    switch (op) {
        case AXHEL_HIP_OP_ADD:  // eltwise-{add,sub}-mod.cpp: vv with operand2, else vs with scalars[l]
        case AXHEL_HIP_OP_SUB:
            if (operand2) {
                eop = op == AXHEL_HIP_OP_ADD ? ELT_ADD_VV : ELT_SUB_VV;
                b = operand2;
            } else if (scalars) {
                eop = op == AXHEL_HIP_OP_ADD ? ELT_ADD_VS : ELT_SUB_VS;
                for (uint32_t l = 0; l < L; ++l) sc.s[l] = scalars[l];
            } else {
                return AXHEL_HIP_INVALID_ARGUMENT;
            }
            break;
        case AXHEL_HIP_OP_MUL:  // eltwise-mul-mod.cpp:81-94: F in {1, 2, 4}
            if (!operand2) return AXHEL_HIP_INVALID_ARGUMENT;
            if (input_mod_factor != 1 && input_mod_factor != 2 && input_mod_factor != 4)
                return AXHEL_HIP_UNSUPPORTED_FACTOR;
            eop = ELT_MUL;
            f = (int)input_mod_factor;
            b = operand2;
            break;
        case AXHEL_HIP_OP_FMA: {  // eltwise-fma-mod.cpp:96-112: F in {1, 2, 4, 8}, addend nullable
            if (!scalars) return AXHEL_HIP_INVALID_ARGUMENT;
            if (input_mod_factor != 1 && input_mod_factor != 2 && input_mod_factor != 4 && input_mod_factor != 8)
                return AXHEL_HIP_UNSUPPORTED_FACTOR;
            eop = operand2 ? ELT_FMA : ELT_FMA_NULL;
            f = (int)input_mod_factor;
            b = operand2;
            for (uint32_t l = 0; l < L; ++l) {  // ReduceInputNative<F>(scalar), eltwise-fma-mod.cpp:32
                const uint64_t q = ctx->plans[l]->q;
                uint64_t s = scalars[l];
                if (f == 8 && s >= (q << 2)) s -= q << 2;
                if (f >= 4 && s >= (q << 1)) s -= q << 1;
                if (f >= 2 && s >= q) s -= q;
                sc.s[l] = s;
            }
            break;
        }
        case AXHEL_HIP_OP_REDUCE: {  // branch order of eltwise-reduce-mod.cpp:32-128
            const uint64_t fi = input_mod_factor, fo = output_mod_factor;
            if (fi == 1 || (fo == 2 && fi <= 2)) {
                if (result == operand1) return AXHEL_HIP_OK;
                AXHEL_HIP_TRY(hipMemcpyAsync(result, operand1, count * L * ctx->n * sizeof(uint64_t),
                                             hipMemcpyDeviceToDevice, st));
                return AXHEL_HIP_OK;
            }
            if (fi == 2 && fo == 1) eop = ELT_RED_2_1;
            else if (fi == 4 && fo == 1) eop = ELT_RED_4_1;
            else if (fi == 4 && fo == 2) eop = ELT_RED_4_2;
            else if (fi == AXHEL_HIP_RNS_ANY_INPUT) eop = ELT_RED_BARRETT;  // in_mod_factor == modulus, every limb
            else return AXHEL_HIP_UNSUPPORTED_FACTOR;
            break;
        }
        default:
            return AXHEL_HIP_INVALID_ARGUMENT;
    }
    const uint64_t nl = count * L;
    std::lock_guard<std::mutex> lk(ctx->mu);
    axhel_hip_rns::StreamTable *tab_buf = nullptr;
    AXHEL_HIP_TRY(stream_table(ctx, st, nl * sizeof(EltLimbDev), &tab_buf));
    EltLimbDev *d = (EltLimbDev *)tab_buf->buf;
    // the table depends on the buffers, the count and the per-limb scalars: a
    // repeated call finds it in place
    const bool same = tab_buf->elt_valid && tab_buf->e_r == result && tab_buf->e_a == operand1 &&
                      tab_buf->e_b == b && tab_buf->e_count == count &&
                      std::memcmp(tab_buf->e_s, sc.s, L * sizeof(uint64_t)) == 0;
    hipError_t e = hipSuccess;
    if (!same) {
        tab_buf->ntt_valid = false;  // the element-wise table overwrites the buffer
        tab_buf->elt_valid = false;
        hipLaunchKernelGGL(rns_fill_elt, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, d, ctx->d_info,
                           result, operand1, b, (uint32_t)count, L, ctx->n, sc);
        e = hipGetLastError();
        if (e == hipSuccess) {
            tab_buf->elt_valid = true;
            tab_buf->e_r = result;
            tab_buf->e_a = operand1;
            tab_buf->e_b = b;
            tab_buf->e_count = count;
            std::memcpy(tab_buf->e_s, sc.s, L * sizeof(uint64_t));
        }
    }
    if (e == hipSuccess) e = launch_elt_batch(eop, f, d, (uint32_t)nl, ctx->n, st);
    AXHEL_HIP_TRY(e);
    return AXHEL_HIP_OK;
}

}  // extern "C"
