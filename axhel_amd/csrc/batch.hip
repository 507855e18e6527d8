// Batched NTT: many limb-transforms (polynomials x RNS limbs, uneven degrees)
// behind one reusable handle whose launches are captured into one hipGraph.
//
// Jobs are grouped by (degree, direction, normalisation); each group's
// descriptors live in device memory, so one kernel launch per pass covers the
// whole group whatever mix of moduli it holds.  Inside a group the jobs are
// ordered by plan, so limbs sharing a twiddle table are adjacent and (through
// the kernels' XCD-contiguous tile order) land on the same XCD's L2.  Groups
// are independent, so the graph forks them over several capture streams and
// the device may run them concurrently.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "../../include/axhel_hip_ubench.h"
#include "ntt_host.hpp"

using axhel_dev::NttJobDev;
using namespace axhel_host;

struct axhel_hip_batch {
    struct Group {
        uint32_t logn;
        bool inverse;
        bool normalize;
        uint32_t count;
        size_t offset;  // first descriptor in d_jobs
        bool vec16;     // all src/dst 16-byte aligned
        bool relax_ok;  // every modulus below 2^60 (relaxed forward schedule allowed)
        // stacked ROW pass: stack table at d_jobs + stack_offset (stack_gpw 0:
        // none), the jobs outside every stack at d_jobs + rest_offset
        size_t stack_offset;
        uint32_t stacks;
        int stack_gpw;
        size_t rest_offset;
        uint32_t rest_count;
    };
    std::vector<Group> groups;
    NttJobDev *d_jobs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> side;  // capture streams for parallel branches
    std::vector<hipEvent_t> events;
};

namespace {

constexpr int kBranches = 4;

void destroy(axhel_hip_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto e : b->events) (void)hipEventDestroy(e);
    for (auto s : b->side) (void)hipStreamDestroy(s);
    (void)hipFree(b->d_jobs);
    delete b;
}

StackTable stack_of(const axhel_hip_batch *b, const axhel_hip_batch::Group &g) {
    StackTable t;
    if (g.stack_gpw) {
        t.jobs = b->d_jobs + g.stack_offset;
        t.count = g.stacks;
        t.gpw = g.stack_gpw;
        t.rest = g.rest_count ? b->d_jobs + g.rest_offset : nullptr;
        t.rest_count = g.rest_count;
    }
    return t;
}

hipError_t launch_group(const axhel_hip_batch *b, const axhel_hip_batch::Group &g, hipStream_t stream,
                        int only_pass = -1) {
    static const NttJobDev none{};
    const StackTable st = stack_of(b, g);
    return launch_ntt(config_for_launch(g.logn, g.count), g.logn, g.inverse, g.normalize, b->d_jobs + g.offset, none,
                      g.count,
                      stream, only_pass, g.vec16, &st, g.relax_ok);
}

hipError_t issue(axhel_hip_batch *b, hipStream_t stream) {
    for (const auto &g : b->groups) {
        hipError_t e = launch_group(b, g, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Capture.  One group: its passes are captured straight on the origin stream
// (a plain chain of kernel nodes).  Several groups: each is forked onto a
// side stream so independent groups become parallel graph branches, joined
// back on the origin stream.
hipError_t capture(axhel_hip_batch *b) {
    hipStream_t origin = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&origin, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    const int nb = b->groups.size() > 1 ? std::min<int>(kBranches, (int)b->groups.size()) : 0;
    for (int i = 0; i < nb; ++i) {
        hipStream_t s;
        hipEvent_t ev;
        if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) break;
        b->side.push_back(s);
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) break;
        b->events.push_back(ev);
    }
    hipEvent_t fork_ev = nullptr;
    if (e == hipSuccess && nb) e = hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t ce = hipSuccess;
        if (nb == 0) {
            ce = launch_group(b, b->groups[0], origin);
        } else {
            ce = hipEventRecord(fork_ev, origin);
            for (int i = 0; i < nb && ce == hipSuccess; ++i) ce = hipStreamWaitEvent(b->side[i], fork_ev, 0);
            for (size_t gi = 0; gi < b->groups.size() && ce == hipSuccess; ++gi) {
                ce = launch_group(b, b->groups[gi], b->side[gi % nb]);
            }
            for (int i = 0; i < nb && ce == hipSuccess; ++i) {
                ce = hipEventRecord(b->events[i], b->side[i]);
                if (ce == hipSuccess) ce = hipStreamWaitEvent(origin, b->events[i], 0);
            }
        }
        e = hipStreamEndCapture(origin, &b->graph);
        if (ce != hipSuccess) e = ce;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    (void)hipStreamDestroy(origin);
    return e;
}

}  // namespace

extern "C" {

axhel_hip_status axhel_hip_ntt_batch_create(const axhel_hip_ntt_job *jobs, size_t count, axhel_hip_batch **batch) {
    if (!batch || (count && !jobs)) return AXHEL_HIP_INVALID_ARGUMENT;
    *batch = nullptr;
    if (count > 0xFFFFFFFFull) return AXHEL_HIP_INVALID_ARGUMENT;
    for (size_t i = 0; i < count; ++i) {
        const auto &j = jobs[i];
        if (!j.plan || !j.result || !j.operand || (j.inverse != 0 && j.inverse != 1))
            return AXHEL_HIP_INVALID_ARGUMENT;
        if (!is_device_pointer(j.result) || !is_device_pointer(j.operand)) return AXHEL_HIP_INVALID_ARGUMENT;
    }
    axhel_hip_batch *b = new (std::nothrow) axhel_hip_batch;
    if (!b) return AXHEL_HIP_NO_MEMORY;

    // group key: (logn, inverse, normalize); inside a group order by plan
    using Key = std::tuple<uint32_t, int, int>;
    std::map<Key, std::vector<size_t>> by;
    for (size_t i = 0; i < count; ++i) {
        const auto &j = jobs[i];
        const bool inv = j.inverse == 1;
        const bool norm = inv ? j.output_mod_factor != 2 : j.output_mod_factor != 4;
        by[Key(j.plan->logn, inv ? 1 : 0, norm ? 1 : 0)].push_back(i);
    }
    std::vector<NttJobDev> host;
    host.reserve(count);
    for (auto &kv : by) {
        auto &idx = kv.second;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t c) {
            return std::less<const void *>()(jobs[a].plan, jobs[c].plan);
        });
        axhel_hip_batch::Group g;
        g.logn = std::get<0>(kv.first);
        g.inverse = std::get<1>(kv.first) != 0;
        g.normalize = std::get<2>(kv.first) != 0;
        g.count = (uint32_t)idx.size();
        g.offset = host.size();
        g.vec16 = true;
        g.relax_ok = true;
        for (size_t i : idx) {
            host.push_back(make_job(jobs[i].plan, jobs[i].operand, jobs[i].result, g.inverse));
            g.vec16 = g.vec16 && aligned16(jobs[i].operand) && aligned16(jobs[i].result);
            g.relax_ok = g.relax_ok && relax_ok_modulus(jobs[i].plan->q);
        }
        g.stack_offset = 0;
        g.stacks = 0;
        g.stack_gpw = 0;
        g.rest_offset = 0;
        g.rest_count = 0;
        b->groups.push_back(g);
    }
    // Stack tables and the jobs left out of them (after all plain
    // descriptors) for groups whose ROW pass can run over same-plan stacks.
    for (auto &g : b->groups) {
        const int gpw = stack_gpw_of(config_for_launch(g.logn, g.count));
        std::vector<NttJobDev> st, rest;
        uint32_t stacks = 0;
        if (gpw && build_stacks(host.data() + g.offset, g.count, gpw, st, stacks, rest)) {
            g.stack_offset = host.size();
            g.stacks = stacks;
            g.stack_gpw = gpw;
            host.insert(host.end(), st.begin(), st.end());
            g.rest_offset = host.size();
            g.rest_count = (uint32_t)rest.size();
            host.insert(host.end(), rest.begin(), rest.end());
        }
    }
    // Largest groups first: they fill the device while smaller ones trail.
    std::stable_sort(b->groups.begin(), b->groups.end(),
                     [](const axhel_hip_batch::Group &x, const axhel_hip_batch::Group &y) {
                         return ((uint64_t)x.count << x.logn) > ((uint64_t)y.count << y.logn);
                     });
    hipError_t e = hipSuccess;
    if (count) {
        e = hipMalloc(&b->d_jobs, host.size() * sizeof(NttJobDev));
        if (e == hipSuccess)
            e = hipMemcpy(b->d_jobs, host.data(), host.size() * sizeof(NttJobDev), hipMemcpyHostToDevice);
    }
    for (const auto &g : b->groups)
        if (e == hipSuccess) e = prepare_config(config_for_launch(g.logn, g.count));
    if (e == hipSuccess && count) e = capture(b);
    if (e != hipSuccess) {
        set_last_error("batch create", e);
        destroy(b);
        return e == hipErrorOutOfMemory ? AXHEL_HIP_NO_MEMORY : AXHEL_HIP_HIP_ERROR;
    }
    *batch = b;
    return AXHEL_HIP_OK;
}

// A batch of one group is a plain chain of its passes: the graph gives it no
// parallel branches, and back-to-back replays of one graph ran 1.5-2.5 %
// slower than the same kernels launched directly (0.629 against 0.614 ms for
// 64 x 24 x 2^16), so launch issues them directly.  Several groups replay
// the graph, whose branches run the groups side by side.
axhel_hip_status axhel_hip_ntt_batch_launch(axhel_hip_batch *batch, void *stream) {
    if (!batch) return AXHEL_HIP_INVALID_ARGUMENT;
    if (batch->groups.size() == 1) {
        AXHEL_HIP_TRY(issue(batch, as_stream(stream)));
        return AXHEL_HIP_OK;
    }
    return axhel_hip_ntt_batch_launch_graph(batch, stream);
}

axhel_hip_status axhel_hip_ntt_batch_launch_graph(axhel_hip_batch *batch, void *stream) {
    if (!batch) return AXHEL_HIP_INVALID_ARGUMENT;
    if (!batch->exec) return AXHEL_HIP_OK;  // empty batch
    AXHEL_HIP_TRY(hipGraphLaunch(batch->exec, as_stream(stream)));
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_ntt_batch_launch_eager(axhel_hip_batch *batch, void *stream) {
    if (!batch) return AXHEL_HIP_INVALID_ARGUMENT;
    AXHEL_HIP_TRY(issue(batch, as_stream(stream)));
    return AXHEL_HIP_OK;
}

int axhel_hip_ntt_batch_kernel_count(const axhel_hip_batch *batch) {
    if (!batch) return 0;
    int k = 0;
    for (const auto &g : batch->groups) k += config_for_launch(g.logn, g.count).npass;
    return k;
}

int axhel_hip_ntt_batch_stacked_groups(const axhel_hip_batch *batch) {
    if (!batch) return 0;
    int k = 0;
    for (const auto &g : batch->groups)
        k += g.stack_gpw != 0 && g.stack_gpw == stack_gpw_of(config_for_launch(g.logn, g.count));
    return k;
}

axhel_hip_status axhel_hip_ntt_batch_launch_kernel(axhel_hip_batch *batch, int index, void *stream) {
    if (!batch || index < 0) return AXHEL_HIP_INVALID_ARGUMENT;
    for (const auto &g : batch->groups) {
        const NttConfig &cfg = config_for_launch(g.logn, g.count);
        if (index < cfg.npass) {
            AXHEL_HIP_TRY(launch_group(batch, g, as_stream(stream), index));
            return AXHEL_HIP_OK;
        }
        index -= cfg.npass;
    }
    return AXHEL_HIP_INVALID_ARGUMENT;
}

axhel_hip_status axhel_hip_ntt_batch_kernel_info(const axhel_hip_batch *batch, int index,
                                                 axhel_hip_ntt_kernel_info *info) {
    if (!batch || !info || index < 0) return AXHEL_HIP_INVALID_ARGUMENT;
    for (const auto &g : batch->groups) {
        const NttConfig &cfg = config_for_launch(g.logn, g.count);
        if (index >= cfg.npass) {
            index -= cfg.npass;
            continue;
        }
        // the choices launch_ntt makes for this pass
        const int p = index;
        const bool last = p == cfg.npass - 1;
        const bool row_pass = cfg.npass == 2 && p == (g.inverse ? 0 : 1);
        const PassKernel &ks = g.inverse ? cfg.inv_stack : cfg.fwd_stack;
        const bool stacked = row_pass && g.stack_gpw && ks.fn && ks.gpw == g.stack_gpw;
        const PassKernel &k = stacked ? ks : (g.inverse ? cfg.inv[p] : cfg.fwd[p]);
        const bool relaxed = !g.inverse && forward_relaxed(cfg, g.normalize, g.relax_ok);
        std::memset(info, 0, sizeof(*info));
        info->degree = 1ull << g.logn;
        info->limbs = g.count;
        info->pass = (uint32_t)p;
        info->passes = (uint32_t)cfg.npass;
        info->log_bits = (uint32_t)k.log_g;
        info->inverse = g.inverse;
        info->normalize = last && g.normalize;
        info->relaxed = relaxed;
        info->stacked = stacked;
        if (!g.inverse)
            info->schedule = !relaxed ? AXHEL_HIP_UBENCH_FWD_EXACT
                             : last   ? AXHEL_HIP_UBENCH_FWD_RELAXED_LAST
                                      : AXHEL_HIP_UBENCH_FWD_RELAXED;
        else
            info->schedule = !last         ? AXHEL_HIP_UBENCH_INV
                             : g.normalize ? AXHEL_HIP_UBENCH_INV_FINAL_NORM
                                           : AXHEL_HIP_UBENCH_INV_FINAL;
        info->butterflies = (info->degree / 2) * info->log_bits * g.count;
        info->bytes = 16ull * info->degree * g.count;
        return AXHEL_HIP_OK;
    }
    return AXHEL_HIP_INVALID_ARGUMENT;
}

axhel_hip_status axhel_hip_ntt_batch_destroy(axhel_hip_batch *batch) {
    destroy(batch);
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_ntt_batch_run(const axhel_hip_ntt_job *jobs, size_t count, void *stream) {
    axhel_hip_batch *b = nullptr;
    axhel_hip_status s = axhel_hip_ntt_batch_create(jobs, count, &b);
    if (s != AXHEL_HIP_OK) return s;
    s = axhel_hip_ntt_batch_launch_eager(b, stream);
    if (s == AXHEL_HIP_OK) s = axhel_hip_stream_synchronize(stream);
    destroy(b);
    return s;
}

}  // extern "C"
