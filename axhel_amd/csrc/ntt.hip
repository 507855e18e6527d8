// NTT plans, twiddle tables, kernel configurations and the plain per-call
// C-ABI entry points (ComputeForward / ComputeInverse drop-ins).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "eltwise_host.hpp"
#include "ntt_host.hpp"
#include "ntt_registry.hpp"
#include "polymul_kernels.hpp"

using axhel_dev::NttJobDev;
using axhel_dev::PassArgs;

namespace axhel_host {
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

uint64_t shoup_quotient(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

}  // namespace
}  // namespace axhel_host

namespace axhel_dev {

// psi^e mod q by square-and-multiply; base < q.  barrett_mul is the exact
// product mod q for operands below q, so every power equals the host's
// (u128)a * b % q chain.
__device__ __forceinline__ uint64_t powmod_dev(uint64_t b, uint64_t e, const Barrett &bc) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = barrett_mul(r, b, bc);
        e >>= 1;
        if (e) b = barrett_mul(b, b, bc);
    }
    return r;
}

// floor(w * 2^64 / q) for w < q < 2^62: restoring long division of w * 2^64
// by q, one quotient bit per step (the remainder stays below 2q < 2^63).
__device__ __forceinline__ uint64_t shoup_quotient_dev(uint64_t w, uint64_t q) {
    uint64_t rem = w, quo = 0;
    for (int b = 0; b < 64; ++b) {
        rem <<= 1;
        quo <<= 1;
        if (rem >= q) {
            rem -= q;
            quo |= 1;
        }
    }
    return quo;
}

// Twiddle tables of ntt-tables.cpp:81-140, one entry per thread:
// fwd[i] = psi^brv(i); inv[0] = {0, 0}, inv[i] = psi^-(brv(i-1)+1), each with
// its Shoup quotient.  The reference walks the bit-reversed index
// incrementally and builds powers sequentially; entry i here is computed
// directly, the same value.  psi and ipsi are reduced mod q.
__global__ void gen_tables_kernel(ulonglong2 *fwd, ulonglong2 *inv, uint32_t logn, uint64_t q, uint64_t psi,
                                  uint64_t ipsi, Barrett bc) {
    const uint64_t n = 1ull << logn;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t w = powmod_dev(psi, __brevll(i) >> (64 - logn), bc);
    fwd[i] = make_ulonglong2(w, shoup_quotient_dev(w, q));
    if (i == 0) {
        inv[0] = make_ulonglong2(0, 0);
    } else {
        const uint64_t wi = powmod_dev(ipsi, (__brevll(i - 1) >> (64 - logn)) + 1, bc);
        inv[i] = make_ulonglong2(wi, shoup_quotient_dev(wi, q));
    }
}

}  // namespace axhel_dev

namespace axhel_host {
namespace {

uint64_t brv_host(uint64_t i, uint32_t logn) {
    uint64_t r = 0;
    for (uint32_t b = 0; b < logn; ++b) r |= ((i >> b) & 1ull) << (logn - 1 - b);
    return r;
}

std::vector<std::vector<NttConfig>> build_registry() {
    Registry r(kMaxLogN + 1);
    add_configs_1_13(r);
    add_configs_14_15(r);
    add_configs_16a(r);
    add_configs_16b(r);
    add_configs_17(r);
    add_configs_18_19(r);
    add_configs_20(r);
    return r;
}

std::once_flag g_reg_once;
std::vector<std::vector<NttConfig>> *g_reg = nullptr;
std::atomic<int> g_selected[kMaxLogN + 1];
// configuration of single-limb calls (the plain per-call path); -1: g_selected
std::atomic<int> g_single[kMaxLogN + 1];
// configuration of launches of 2 limbs up to kLatencyWords words; -1: g_single
std::atomic<int> g_small[kMaxLogN + 1];
// the tuning table's choices (written once, before any selection), which
// config_select(-1) restores
int g_tuned[kMaxLogN + 1];
int g_tuned_single[kMaxLogN + 1];
int g_tuned_small[kMaxLogN + 1];

// Tuned configuration table written by axhel_amd/autotune.py: lines of
// "log2(N) index [single-call index [small-launch index]]" (the index
// batches use, optionally the one single-limb calls use, whose latency
// favours other shapes, and the one launches of a few limbs use).  Read
// from $AXHEL_HIP_TUNING, else from
// ntt_tuning_gfx950.txt in the package directory above this library.
void load_tuning(const std::vector<std::vector<NttConfig>> &reg) {
    std::string path;
    if (const char *env = std::getenv("AXHEL_HIP_TUNING")) {
        path = env;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<const void *>(&load_tuning), &info) && info.dli_fname) {
            std::string so = info.dli_fname;
            const size_t a = so.find_last_of('/');
            const std::string libdir = a == std::string::npos ? "." : so.substr(0, a);
            const size_t b = libdir.find_last_of('/');
            const std::string pkg = b == std::string::npos ? "." : libdir.substr(0, b);
            path = pkg + "/ntt_tuning_gfx950.txt";
        }
    }
    if (path.empty()) return;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        int logn = -1, idx = -1, single = -1, small = -1;
        const int got = std::sscanf(line, "%d %d %d %d", &logn, &idx, &single, &small);
        if (got >= 2 && logn >= 1 && logn <= (int)kMaxLogN && idx >= 0 && idx < (int)reg[logn].size()) {
            g_tuned[logn] = idx;
            if (got >= 3 && single >= 0 && single < (int)reg[logn].size()) g_tuned_single[logn] = single;
            if (got == 4 && small >= 0 && small < (int)reg[logn].size()) g_tuned_small[logn] = small;
        }
    }
    std::fclose(f);
}

const std::vector<std::vector<NttConfig>> &registry() {
    std::call_once(g_reg_once, [] {
        g_reg = new std::vector<std::vector<NttConfig>>(build_registry());
        for (uint32_t l = 0; l <= kMaxLogN; ++l) {
            g_tuned[l] = 0;
            g_tuned_single[l] = -1;
            g_tuned_small[l] = -1;
        }
        load_tuning(*g_reg);
        for (uint32_t l = 0; l <= kMaxLogN; ++l) {
            g_selected[l].store(g_tuned[l]);
            g_single[l].store(g_tuned_single[l]);
            g_small[l].store(g_tuned_small[l]);
        }
    });
    return *g_reg;
}

// Kernels using more than the default 64 KiB of dynamic LDS need the limit
// raised once per kernel and device.
std::mutex g_attr_mu;
std::vector<std::pair<const void *, int>> g_attr_done;

hipError_t ensure_fn_lds(const void *fn, size_t bytes, int dev) {
    std::lock_guard<std::mutex> lk(g_attr_mu);
    for (auto &d : g_attr_done)
        if (d.first == fn && d.second == dev) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) g_attr_done.emplace_back(fn, dev);
    return e;
}

hipError_t ensure_lds(const PassKernel &k) {
    if (k.lds_bytes <= 65536 && k.lds_bytes_relaxed <= 65536) return hipSuccess;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess && k.lds_bytes > 65536) e = ensure_fn_lds(k.fn, k.lds_bytes, dev);
    if (e == hipSuccess && k.fn_relaxed && k.lds_bytes_relaxed > 65536)
        e = ensure_fn_lds(k.fn_relaxed, k.lds_bytes_relaxed, dev);
    return e;
}

}  // namespace

const std::vector<NttConfig> &configs_for(uint32_t logn) {
    static const std::vector<NttConfig> empty;
    if (logn == 0 || logn > kMaxLogN) return empty;
    return registry()[logn];
}

const NttConfig &config_for(uint32_t logn) {
    const auto &v = configs_for(logn);
    int i = g_selected[logn].load();
    if (i < 0 || i >= (int)v.size()) i = 0;
    return v[i];
}

const NttConfig &config_for_single(uint32_t logn) {
    const auto &v = configs_for(logn);
    const int i = g_single[logn].load();
    if (i >= 0 && i < (int)v.size()) return v[i];
    return config_for(logn);
}

const NttConfig &config_for_small(uint32_t logn) {
    const auto &v = configs_for(logn);
    const int i = g_small[logn].load();
    if (i >= 0 && i < (int)v.size()) return v[i];
    return config_for_single(logn);
}

const NttConfig &config_for_launch(uint32_t logn, uint64_t count) {
    if (count <= 1) return config_for_single(logn);
    return count <= (kLatencyWords >> logn) ? config_for_small(logn) : config_for(logn);
}

hipError_t prepare_config(const NttConfig &cfg) {
    for (int p = 0; p < cfg.npass; ++p) {
        hipError_t e = ensure_lds(cfg.fwd[p]);
        if (e == hipSuccess) e = ensure_lds(cfg.inv[p]);
        if (e != hipSuccess) return e;
    }
    if (cfg.fwd_stack.fn) {
        hipError_t e = ensure_lds(cfg.fwd_stack);
        if (e == hipSuccess) e = ensure_lds(cfg.inv_stack);
        if (e == hipSuccess && cfg.fused_stack.fn) e = ensure_lds(cfg.fused_stack);
        if (e != hipSuccess) return e;
    }
    if (cfg.fused.fn) return ensure_lds(cfg.fused);
    return hipSuccess;
}

int stack_gpw(uint32_t logn) {
    if (configs_for(logn).empty()) return 0;
    return stack_gpw_of(config_for(logn));
}

bool build_stacks(const NttJobDev *jobs, uint32_t count, int gpw, std::vector<NttJobDev> &out, uint32_t &stacks,
                  std::vector<NttJobDev> &rest) {
    out.clear();
    rest.clear();
    stacks = 0;
    if (gpw <= 0 || count == 0) return false;
    const uint32_t g = (uint32_t)gpw;
    for (uint32_t i = 0; i < count;) {
        uint32_t e = i + 1;
        while (e < count && jobs[e].tw == jobs[i].tw && jobs[e].q == jobs[i].q) ++e;
        const uint32_t run = e - i, tail = run % g;
        // a partial stack with at most one empty slot in eight is padded
        const bool pad = tail && (g - tail) * 8 <= g;
        const uint32_t stacked = pad ? run : run - tail;
        for (uint32_t k = i; k < i + stacked; ++k) out.push_back(jobs[k]);
        if (pad) {
            NttJobDev empty = jobs[i];
            empty.src = nullptr;
            empty.dst = nullptr;
            while (out.size() % g) out.push_back(empty);
        }
        for (uint32_t k = i + stacked; k < e; ++k) rest.push_back(jobs[k]);
        i = e;
    }
    if (out.empty()) {
        rest.clear();
        return false;
    }
    stacks = (uint32_t)(out.size() / g);
    return true;
}

NttJobDev make_job(const axhel_hip_plan *p, const uint64_t *src, uint64_t *dst, bool inverse) {
    NttJobDev j;
    j.src = src;
    j.dst = dst;
    j.tw = inverse ? p->d_inv : p->d_fwd;
    j.q = p->q;
    j.ninv = p->ninv;
    j.ninv_q = p->ninv_q;
    j.sroot = p->sroot;
    j.sroot_q = p->sroot_q;
    return j;
}

bool forward_relaxed(const NttConfig &cfg, bool normalize, bool relax_ok) {
    // every pass of the transform on one schedule: relaxed only when each
    // forward kernel that may run has its relaxed twin
    bool relaxed = normalize && relax_ok;
    for (int p = 0; p < cfg.npass && relaxed; ++p) relaxed = cfg.fwd[p].fn_relaxed != nullptr;
    if (relaxed && cfg.fwd_stack.fn && !cfg.fwd_stack.fn_relaxed) relaxed = false;
    return relaxed;
}

hipError_t launch_ntt(const NttConfig &cfg, uint32_t logn, bool inverse, bool normalize, const NttJobDev *jobs,
                      const NttJobDev &inline_job, uint32_t count, hipStream_t stream, int only_pass, bool vec16,
                      const StackTable *stacks, bool relax_ok) {
    if (count == 0) return hipSuccess;
    const uint64_t n = 1ull << logn;
    const bool relaxed = !inverse && forward_relaxed(cfg, normalize, relax_ok);
    for (int p = 0; p < cfg.npass; ++p) {
        if (only_pass >= 0 && p != only_pass) continue;
        const bool row_pass = cfg.npass == 2 && p == (inverse ? 0 : 1);
        const PassKernel &ks = inverse ? cfg.inv_stack : cfg.fwd_stack;
        const bool stacked = row_pass && stacks && stacks->jobs && stacks->count && ks.fn && ks.gpw == stacks->gpw;
        const PassKernel &k = stacked ? ks : (inverse ? cfg.inv[p] : cfg.fwd[p]);
        hipError_t e = ensure_lds(k);
        if (e != hipSuccess) return e;
        PassArgs a;
        std::memset(&a, 0, sizeof(a));
        a.jobs = jobs;
        a.inline_job = inline_job;
        a.logn = logn;
        a.num_limbs = count;
        a.groups_per_limb = (uint32_t)(n >> k.log_g);
        a.first_pass = p == 0;
        a.normalize = (p == cfg.npass - 1) && normalize;
        a.final_inv = inverse && (p == cfg.npass - 1);
        a.vec16 = vec16 ? 1u : 0u;
        uint64_t blocks;
        if (stacked) {
            a.jobs = stacks->jobs;
            a.num_limbs = stacks->count;
            blocks = (uint64_t)stacks->count * a.groups_per_limb;
        } else if (k.multi) {
            blocks = (count + k.gpw - 1) / k.gpw;
        } else {
            blocks = (uint64_t)count * a.groups_per_limb / k.gpw;
        }
        // the kernels index groups with 32-bit arithmetic (tile * GPW + g)
        if (blocks == 0 || blocks * (uint64_t)k.gpw > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
        void *args[] = {&a};
        e = hipLaunchKernel(relaxed ? k.fn_relaxed : k.fn, dim3((unsigned)blocks), dim3(k.threads), args,
                            relaxed ? k.lds_bytes_relaxed : k.lds_bytes, stream);
        if (e != hipSuccess) return e;
        if (stacked && stacks->rest_count) {
            // the jobs outside every stack: the per-limb ROW kernel of the same pass
            const PassKernel &kr = inverse ? cfg.inv[p] : cfg.fwd[p];
            if (kr.multi) return hipErrorInvalidConfiguration;
            if ((e = ensure_lds(kr)) != hipSuccess) return e;
            PassArgs r = a;
            r.jobs = stacks->rest;
            r.num_limbs = stacks->rest_count;
            r.groups_per_limb = (uint32_t)(n >> kr.log_g);
            const uint64_t rblocks = (uint64_t)stacks->rest_count * r.groups_per_limb / kr.gpw;
            if (rblocks == 0 || rblocks * (uint64_t)kr.gpw > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
            void *rargs[] = {&r};
            e = hipLaunchKernel(relaxed ? kr.fn_relaxed : kr.fn, dim3((unsigned)rblocks), dim3(kr.threads), rargs,
                                relaxed ? kr.lds_bytes_relaxed : kr.lds_bytes, stream);
            if (e != hipSuccess) return e;
        }
    }
    return hipGetLastError();
}

}  // namespace axhel_host

// ------------------------------------------------------------------ C ABI

using namespace axhel_host;

namespace {

// The scaled root of the final inverse stage from wl = inv_root_powers[N-1]
// (ntt-native.cpp:109-113, once per table instead of once per call).
void set_final_root(axhel_hip_plan *p, uint64_t wl) {
    uint64_t s = wl * p->ninv - (uint64_t)(((u128)wl * p->ninv_q) >> 64) * p->q;
    if (s >= p->q) s -= p->q;
    p->sroot = s;
    p->sroot_q = shoup_quotient(s, p->q);
}

// Finishes a plan whose ninv, ninv_q are set: the final inverse stage's
// scaled root from wl = inv_root_powers[N-1].operand, then the device
// tables: uploaded from h_fwd/h_inv when the caller supplied them (external
// tables), else generated on the device from psi.
axhel_hip_status finish_plan(axhel_hip_plan *p, uint64_t wl, uint64_t psi, uint64_t ipsi, axhel_hip_plan **out) {
    set_final_root(p, wl);
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc(&p->d_fwd, p->n * sizeof(ulonglong2));
    if (e == hipSuccess) e = hipMalloc(&p->d_inv, p->n * sizeof(ulonglong2));
    if (!p->h_fwd.empty()) {
        if (e == hipSuccess)
            e = hipMemcpy(p->d_fwd, p->h_fwd.data(), p->n * sizeof(ulonglong2), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(p->d_inv, p->h_inv.data(), p->n * sizeof(ulonglong2), hipMemcpyHostToDevice);
    } else if (e == hipSuccess) {
        const unsigned threads = 256;
        const unsigned blocks = (unsigned)((p->n + threads - 1) / threads);
        hipLaunchKernelGGL(axhel_dev::gen_tables_kernel, dim3(blocks), dim3(threads), 0, nullptr, p->d_fwd, p->d_inv,
                           p->logn, p->q, psi, ipsi, barrett_for(p->q));
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    }
    if (e != hipSuccess) {
        set_last_error("plan upload", e);
        (void)hipFree(p->d_fwd);
        (void)hipFree(p->d_inv);
        delete p;
        return e == hipErrorOutOfMemory ? AXHEL_HIP_NO_MEMORY : AXHEL_HIP_HIP_ERROR;
    }
    *out = p;
    return AXHEL_HIP_OK;
}

bool valid_degree(uint64_t degree, uint32_t *logn) {
    if (degree < 2 || (degree & (degree - 1)) != 0) return false;
    uint32_t l = 0;
    while ((1ull << l) < degree) ++l;
    if (l > kMaxLogN) return false;
    *logn = l;
    return true;
}

// A plan on borrowed tables reads them as they are at each call, as the
// reference does: when the table this call uses differs from what the
// device holds, every transform issued so far is let finish (they read the
// old one) and the table is reloaded.  The comparison reads the 16 N bytes
// that the reference's own transform reads from the table on every call.
// Called with the plan's ext_mu held; the plan is shared as const, and its
// device copy is the one part a call may rewrite.
hipError_t reload_external(const axhel_hip_plan *cp, bool inverse) {
    axhel_hip_plan *p = const_cast<axhel_hip_plan *>(cp);
    const uint64_t *src = inverse ? p->ext_inv : p->ext_fwd;
    std::vector<uint64_t> &held = inverse ? p->h_inv : p->h_fwd;
    const size_t bytes = 2 * p->n * sizeof(uint64_t);
    if (std::memcmp(src, held.data(), bytes) == 0) return hipSuccess;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    std::memcpy(held.data(), src, bytes);
    e = hipMemcpy(inverse ? p->d_inv : p->d_fwd, held.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && inverse) set_final_root(p, held[2 * (p->n - 1)]);
    return e;
}

axhel_hip_status run_one(const axhel_hip_plan *plan, uint64_t *result, const uint64_t *operand, bool inverse,
                         bool normalize, void *stream) {
    if (!plan || !result || !operand) return AXHEL_HIP_INVALID_ARGUMENT;
    const hipStream_t st = as_stream(stream);
    Staging stg(st);
    const uint64_t *dop = nullptr;
    uint64_t *dres = nullptr;
    AXHEL_HIP_TRY(stg.input(operand, plan->n, &dop));
    AXHEL_HIP_TRY(stg.output(result, plan->n, &dres));
    // A plan on borrowed tables holds its lock from the table check until its
    // kernels are queued, so a reload (which first waits for the device)
    // never rewrites a table that a queued transform has still to read.
    std::unique_lock<std::mutex> lk;
    if (plan->ext_fwd) {
        lk = std::unique_lock<std::mutex>(plan->ext_mu);
        AXHEL_HIP_TRY(reload_external(plan, inverse));
    }
    const NttJobDev job = make_job(plan, dop, dres, inverse);
    const bool vec16 = aligned16(dop) && aligned16(dres);
    AXHEL_HIP_TRY(launch_ntt(config_for_single(plan->logn), plan->logn, inverse, normalize, nullptr, job, 1, st, -1,
                             vec16, nullptr, relax_ok_modulus(plan->q)));
    if (lk.owns_lock()) lk.unlock();
    AXHEL_HIP_TRY(stg.finish());
    return AXHEL_HIP_OK;
}

}  // namespace

extern "C" {

axhel_hip_status axhel_hip_ntt_plan_create(uint64_t degree, uint64_t modulus, uint64_t root_of_unity,
                                           axhel_hip_plan **plan) {
    if (!plan) return AXHEL_HIP_INVALID_ARGUMENT;
    *plan = nullptr;
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn) || modulus < 2 || modulus >= (1ull << 62)) return AXHEL_HIP_INVALID_ARGUMENT;
    if (axhel_hip_device_count() == 0) return AXHEL_HIP_NO_DEVICE;
    axhel_hip_plan *p = new (std::nothrow) axhel_hip_plan;
    if (!p) return AXHEL_HIP_NO_MEMORY;
    p->n = degree;
    p->logn = logn;
    p->q = modulus;
    // scalars on the host: psi^-1 = psi^(q-2), n^-1 = n^(q-2) (ntt-tables.cpp
    // conventions), and inv_root_powers[N-1] for the final-stage root; the
    // tables themselves are generated on the device
    const uint64_t psi = root_of_unity % modulus;
    const uint64_t ipsi = powmod(psi, modulus - 2, modulus);
    p->ninv = powmod(degree % modulus, modulus - 2, modulus);
    p->ninv_q = shoup_quotient(p->ninv, modulus);
    const uint64_t wl = degree > 1 ? powmod(ipsi, brv_host(degree - 2, logn) + 1, modulus) : 0;
    return finish_plan(p, wl, psi, ipsi, plan);
}

axhel_hip_status axhel_hip_ntt_plan_create_external(uint64_t degree, uint64_t modulus, const uint64_t *root_powers,
                                                    const uint64_t *inv_root_powers, uint64_t inv_degree_operand,
                                                    uint64_t inv_degree_quotient, axhel_hip_plan **plan) {
    if (!plan) return AXHEL_HIP_INVALID_ARGUMENT;
    *plan = nullptr;
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn) || modulus < 2 || modulus >= (1ull << 62) || !root_powers || !inv_root_powers)
        return AXHEL_HIP_INVALID_ARGUMENT;
    if (axhel_hip_device_count() == 0) return AXHEL_HIP_NO_DEVICE;
    axhel_hip_plan *p = new (std::nothrow) axhel_hip_plan;
    if (!p) return AXHEL_HIP_NO_MEMORY;
    p->n = degree;
    p->logn = logn;
    p->q = modulus;
    p->h_fwd.assign(root_powers, root_powers + 2 * degree);
    p->h_inv.assign(inv_root_powers, inv_root_powers + 2 * degree);
    p->ext_fwd = root_powers;
    p->ext_inv = inv_root_powers;
    p->ninv = inv_degree_operand;
    p->ninv_q = inv_degree_quotient;
    return finish_plan(p, inv_root_powers[2 * (degree - 1)], 0, 0, plan);
}

axhel_hip_status axhel_hip_ntt_plan_destroy(axhel_hip_plan *plan) {
    if (!plan) return AXHEL_HIP_OK;
    (void)hipFree(plan->d_fwd);
    (void)hipFree(plan->d_inv);
    delete plan;
    return AXHEL_HIP_OK;
}

uint64_t axhel_hip_ntt_plan_degree(const axhel_hip_plan *plan) { return plan ? plan->n : 0; }
uint64_t axhel_hip_ntt_plan_modulus(const axhel_hip_plan *plan) { return plan ? plan->q : 0; }

axhel_hip_status axhel_hip_ntt_plan_tables(const axhel_hip_plan *plan, uint64_t *root_powers,
                                           uint64_t *inv_root_powers, uint64_t *inv_degree) {
    if (!plan) return AXHEL_HIP_INVALID_ARGUMENT;
    // the device tables are the plan's (host copies exist only for external tables)
    if (root_powers) AXHEL_HIP_TRY(hipMemcpy(root_powers, plan->d_fwd, plan->n * sizeof(ulonglong2), hipMemcpyDeviceToHost));
    if (inv_root_powers)
        AXHEL_HIP_TRY(hipMemcpy(inv_root_powers, plan->d_inv, plan->n * sizeof(ulonglong2), hipMemcpyDeviceToHost));
    if (inv_degree) {
        inv_degree[0] = plan->ninv;
        inv_degree[1] = plan->ninv_q;
    }
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_ntt_forward(const axhel_hip_plan *plan, uint64_t *result, const uint64_t *operand,
                                       uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream) {
    (void)input_mod_factor;  // ignored, as in ntt.cpp:173-186
    return run_one(plan, result, operand, false, output_mod_factor != 4, stream);
}

axhel_hip_status axhel_hip_ntt_inverse(const axhel_hip_plan *plan, uint64_t *result, const uint64_t *operand,
                                       uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream) {
    (void)input_mod_factor;  // ignored, as in ntt.cpp:192-204
    return run_one(plan, result, operand, true, output_mod_factor != 2, stream);
}

int axhel_hip_ntt_config_count(uint64_t degree) {
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn)) return 0;
    return (int)configs_for(logn).size();
}

axhel_hip_status axhel_hip_ntt_config_get(uint64_t degree, int index, axhel_hip_ntt_config *config) {
    uint32_t logn = 0;
    if (!config || !valid_degree(degree, &logn)) return AXHEL_HIP_INVALID_ARGUMENT;
    const auto &v = configs_for(logn);
    if (index < 0 || index >= (int)v.size()) return AXHEL_HIP_INVALID_ARGUMENT;
    *config = v[index].pub;
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_ntt_config_select(uint64_t degree, int index) {
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn)) return AXHEL_HIP_INVALID_ARGUMENT;
    const auto &v = configs_for(logn);
    if (index < -1 || index >= (int)v.size()) return AXHEL_HIP_INVALID_ARGUMENT;
    if (index < 0) {  // -1: back to the tuning table's choices (the first configuration without one)
        g_selected[logn].store(g_tuned[logn]);
        g_single[logn].store(g_tuned_single[logn]);
        g_small[logn].store(g_tuned_small[logn]);
    } else {
        g_selected[logn].store(index);
        g_single[logn].store(-1);  // single calls and small launches follow an explicit selection
        g_small[logn].store(-1);
    }
    return AXHEL_HIP_OK;
}

int axhel_hip_ntt_config_selected(uint64_t degree) {
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn)) return -1;
    (void)registry();
    return g_selected[logn].load();
}

int axhel_hip_ntt_config_single(uint64_t degree) {
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn)) return -1;
    (void)registry();
    const int i = g_single[logn].load();
    return i >= 0 ? i : g_selected[logn].load();
}

int axhel_hip_ntt_config_small(uint64_t degree) {
    uint32_t logn = 0;
    if (!valid_degree(degree, &logn)) return -1;
    (void)registry();
    const int i = g_small[logn].load();
    return i >= 0 ? i : axhel_hip_ntt_config_single(degree);
}

}  // extern "C"
