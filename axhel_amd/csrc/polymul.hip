// Batched RNS polynomial product: per job (one limb),
//   result = InverseNTT(FMAMod(MulMod(NTT(a), NTT(b), 1), scalar, addend, 1))
// either fused (forward COL passes of a and b -> one fused ROW kernel ->
// inverse COL pass: three launches, four HBM round trips of a limb) or as the
// plain five-call sequence of batched kernels (the reference's call order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "eltwise_host.hpp"
#include "ntt_host.hpp"
#include "polymul_kernels.hpp"

using axhel_dev::NttJobDev;
using axhel_dev::PolyArgs;
using axhel_dev::PolyJobDev;
using namespace axhel_host;

struct axhel_hip_polymul {
    struct Group {
        uint32_t logn;
        uint32_t count;
        size_t off_col;   // 2*count NttJobDev (a -> result, b -> scratch)
        size_t off_inv;   // count NttJobDev (result in place)
        size_t off_fused; // count PolyJobDev
        size_t off_elt_mul, off_elt_fma;  // count EltLimbDev each, unfused path
        bool vec16;
        bool has_addend;
        bool relax_ok;  // every modulus below 2^60: the forward transforms may run relaxed
        // fused path over stacks of same-plan jobs: stack table at off_fstack
        // (fstack_gpw 0: none)
        size_t off_fstack;
        uint32_t fstacks;
        int fstack_gpw;
        // the configuration the descriptors were built for (their a1/b1
        // inputs depend on its pass count): launches always use it, whatever
        // is selected for the degree later
        int cfg_index;
    };
    std::vector<Group> groups;
    char *d_desc = nullptr;
    bool fused = true;
};

namespace {

template <typename T>
size_t put(std::vector<char> &buf, const std::vector<T> &v) {
    const size_t off = (buf.size() + 63) & ~(size_t)63;
    buf.resize(off + v.size() * sizeof(T));
    std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// Stacks of gpw same-plan product jobs (runs of equal modulus and tables in
// `jobs`, which are ordered by plan), padded with empty slots (out null);
// false when padding would leave more than one slot in eight empty.
bool stack_poly_jobs(const std::vector<PolyJobDev> &jobs, int gpw, std::vector<PolyJobDev> &out, uint32_t &stacks) {
    out.clear();
    stacks = 0;
    const size_t count = jobs.size();
    if (gpw <= 0 || count == 0) return false;
    auto same = [&](size_t x, size_t y) {
        return jobs[x].twf == jobs[y].twf && jobs[x].twi == jobs[y].twi && jobs[x].q == jobs[y].q;
    };
    uint64_t slots = 0;
    for (size_t i = 0; i < count;) {
        size_t e = i + 1;
        while (e < count && same(e, i)) ++e;
        slots += (e - i + gpw - 1) / gpw * gpw;
        i = e;
    }
    if ((slots - count) * 8 > slots) return false;
    for (size_t i = 0; i < count;) {
        size_t e = i + 1;
        while (e < count && same(e, i)) ++e;
        for (size_t k = i; k < e; ++k) out.push_back(jobs[k]);
        PolyJobDev empty = jobs[i];
        empty.out = nullptr;
        while (out.size() % gpw) out.push_back(empty);
        i = e;
    }
    stacks = (uint32_t)(out.size() / gpw);
    return true;
}

const NttConfig &group_config(const axhel_hip_polymul::Group &g) { return configs_for(g.logn)[g.cfg_index]; }

hipError_t issue(axhel_hip_polymul *pm, hipStream_t st) {
    static const NttJobDev none{};
    for (const auto &g : pm->groups) {
        const NttConfig &cfg = group_config(g);
        hipError_t e = prepare_config(cfg);
        if (e != hipSuccess) return e;
        const NttJobDev *col = (const NttJobDev *)(pm->d_desc + g.off_col);
        const NttJobDev *inv = (const NttJobDev *)(pm->d_desc + g.off_inv);
        if (pm->fused && cfg.fused.fn) {
            const bool stacked = g.fstack_gpw && cfg.fused_stack.fn && cfg.fused_stack.gpw == g.fstack_gpw;
            const PassKernel &fk = stacked ? cfg.fused_stack : cfg.fused;
            // the forward COL pass and the fused kernel's forward stages on one
            // schedule (the fused kernel normalizes NTT(a) and NTT(b))
            const bool rlx = forward_relaxed(cfg, true, g.relax_ok) && fk.fn_relaxed;
            if (cfg.npass == 2) {
                // forward COL pass of every a and b (normalize applies to the last pass only, so
                // this pass is the same; it selects the schedule)
                if ((e = launch_ntt(cfg, g.logn, false, true, col, none, 2 * g.count, st, 0, g.vec16, nullptr,
                                    rlx)))
                    return e;
            }
            PolyArgs a;
            std::memset(&a, 0, sizeof(a));
            a.jobs = (const PolyJobDev *)(pm->d_desc + g.off_fused);
            a.logn = g.logn;
            a.num_limbs = g.count;
            a.groups_per_limb = (uint32_t)((1ull << g.logn) >> cfg.fused.log_g);
            a.single = cfg.npass == 1;
            a.vec16 = g.vec16;
            uint64_t blocks = (uint64_t)g.count * a.groups_per_limb / cfg.fused.gpw;
            if (stacked) {
                a.jobs = (const PolyJobDev *)(pm->d_desc + g.off_fstack);
                a.num_limbs = g.fstacks;
                blocks = (uint64_t)g.fstacks * a.groups_per_limb;
            }
            if (blocks == 0 || blocks * (uint64_t)fk.gpw > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
            void *args[] = {&a};
            if ((e = hipLaunchKernel(rlx ? fk.fn_relaxed : fk.fn, dim3((unsigned)blocks), dim3(fk.threads), args,
                                     rlx ? fk.lds_bytes_relaxed : fk.lds_bytes, st)))
                return e;
            if (cfg.npass == 2) {
                if ((e = launch_ntt(cfg, g.logn, true, true, inv, none, g.count, st, 1, g.vec16))) return e;
            }
        } else {
            // ComputeForward(a) -> result, ComputeForward(b) -> scratch: one launch per pass for both
            if ((e = launch_ntt(cfg, g.logn, false, true, col, none, 2 * g.count, st, -1, g.vec16, nullptr,
                                g.relax_ok)))
                return e;
            const uint64_t n = 1ull << g.logn;
            if ((e = launch_elt_batch(ELT_MUL, 1, (const EltLimbDev *)(pm->d_desc + g.off_elt_mul), g.count, n, st)))
                return e;
            if ((e = launch_elt_batch(g.has_addend ? ELT_FMA : ELT_FMA_NULL, 1,
                                      (const EltLimbDev *)(pm->d_desc + g.off_elt_fma), g.count, n, st)))
                return e;
            if ((e = launch_ntt(cfg, g.logn, true, true, inv, none, g.count, st, -1, g.vec16))) return e;
        }
    }
    return hipGetLastError();
}

}  // namespace

extern "C" {

axhel_hip_status axhel_hip_polymul_create(const axhel_hip_polymul_job *jobs, size_t count, int fused,
                                          axhel_hip_polymul **out) {
    if (!out || (count && !jobs) || count > 0x7FFFFFFF) return AXHEL_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    for (size_t i = 0; i < count; ++i) {
        const auto &j = jobs[i];
        if (!j.plan || !j.result || !j.a || !j.b || !j.scratch) return AXHEL_HIP_INVALID_ARGUMENT;
        if (!is_device_pointer(j.result) || !is_device_pointer(j.a) || !is_device_pointer(j.b) ||
            !is_device_pointer(j.scratch) || (j.addend && !is_device_pointer(j.addend)))
            return AXHEL_HIP_INVALID_ARGUMENT;
        if (j.plan->q < 2) return AXHEL_HIP_INVALID_ARGUMENT;
    }
    axhel_hip_polymul *pm = new (std::nothrow) axhel_hip_polymul;
    if (!pm) return AXHEL_HIP_NO_MEMORY;
    pm->fused = fused != 0;
    // group by (degree, has-addend); order by plan inside a group
    std::map<std::pair<uint32_t, int>, std::vector<size_t>> by;
    for (size_t i = 0; i < count; ++i) by[{jobs[i].plan->logn, jobs[i].addend ? 1 : 0}].push_back(i);
    std::vector<char> buf;
    for (auto &kv : by) {
        auto &idx = kv.second;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
            return std::less<const void *>()(jobs[x].plan, jobs[y].plan);
        });
        axhel_hip_polymul::Group g{};
        g.logn = kv.first.first;
        const NttConfig &cfg = config_for(g.logn);
        g.cfg_index = (int)(&cfg - configs_for(g.logn).data());
        g.count = (uint32_t)idx.size();
        g.has_addend = kv.first.second != 0;
        g.vec16 = true;
        g.relax_ok = true;
        std::vector<NttJobDev> col, inv;
        std::vector<PolyJobDev> fz;
        std::vector<EltLimbDev> mul, fma;
        for (size_t i : idx) {
            const auto &j = jobs[i];
            const axhel_hip_plan *p = j.plan;
            g.vec16 = g.vec16 && aligned16(j.a) && aligned16(j.b) && aligned16(j.result) && aligned16(j.scratch) &&
                      (!j.addend || aligned16(j.addend));
            g.relax_ok = g.relax_ok && relax_ok_modulus(p->q);
            col.push_back(make_job(p, j.a, j.result, false));
        }
        for (size_t i : idx) col.push_back(make_job(jobs[i].plan, jobs[i].b, jobs[i].scratch, false));
        for (size_t i : idx) {
            const auto &j = jobs[i];
            inv.push_back(make_job(j.plan, j.result, j.result, true));
            PolyJobDev d{};
            d.a1 = cfg.npass == 2 ? j.result : j.a;
            d.b1 = cfg.npass == 2 ? j.scratch : j.b;
            d.c = j.addend;
            d.out = j.result;
            d.twf = j.plan->d_fwd;
            d.twi = j.plan->d_inv;
            d.q = j.plan->q;
            d.bc = barrett_for(j.plan->q);
            d.scalar = j.scalar;
            d.ninv = j.plan->ninv;
            d.ninv_q = j.plan->ninv_q;
            d.sroot = j.plan->sroot;
            d.sroot_q = j.plan->sroot_q;
            fz.push_back(d);
            EltLimbDev m{};
            m.r = j.result;
            m.a = j.result;
            m.b = j.scratch;
            m.bc = d.bc;
            mul.push_back(m);
            EltLimbDev f{};
            f.r = j.result;
            f.a = j.result;
            f.b = j.addend;
            f.s = j.scalar;  // ReduceInputNative<1>: unchanged
            f.bc = d.bc;
            fma.push_back(f);
        }
        g.off_col = put(buf, col);
        g.off_inv = put(buf, inv);
        g.off_fused = put(buf, fz);
        g.off_elt_mul = put(buf, mul);
        g.off_elt_fma = put(buf, fma);
        g.off_fstack = 0;
        g.fstacks = 0;
        g.fstack_gpw = 0;
        std::vector<PolyJobDev> fst;
        if (cfg.npass == 2 && cfg.fused_stack.fn && stack_poly_jobs(fz, cfg.fused_stack.gpw, fst, g.fstacks)) {
            g.off_fstack = put(buf, fst);
            g.fstack_gpw = cfg.fused_stack.gpw;
        }
        pm->groups.push_back(g);
    }
    hipError_t e = hipSuccess;
    if (!buf.empty()) {
        e = hipMalloc(&pm->d_desc, buf.size());
        if (e == hipSuccess) e = hipMemcpy(pm->d_desc, buf.data(), buf.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        set_last_error("polymul create", e);
        (void)hipFree(pm->d_desc);
        delete pm;
        return e == hipErrorOutOfMemory ? AXHEL_HIP_NO_MEMORY : AXHEL_HIP_HIP_ERROR;
    }
    *out = pm;
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_polymul_launch(axhel_hip_polymul *pm, void *stream) {
    if (!pm) return AXHEL_HIP_INVALID_ARGUMENT;
    AXHEL_HIP_TRY(issue(pm, as_stream(stream)));
    return AXHEL_HIP_OK;
}

int axhel_hip_polymul_stacked_groups(const axhel_hip_polymul *pm) {
    if (!pm) return 0;
    int k = 0;
    for (const auto &g : pm->groups) {
        const NttConfig &cfg = group_config(g);
        k += pm->fused && g.fstack_gpw != 0 && cfg.fused_stack.fn && cfg.fused_stack.gpw == g.fstack_gpw;
    }
    return k;
}

axhel_hip_status axhel_hip_polymul_destroy(axhel_hip_polymul *pm) {
    if (!pm) return AXHEL_HIP_OK;
    (void)hipFree(pm->d_desc);
    delete pm;
    return AXHEL_HIP_OK;
}

}  // extern "C"
