// Serving loop: continuous batching of NTT / element-wise requests.
//
// A scheduler thread owns everything device-side.  Requests arrive from any
// thread (submit), wait in FIFO order, and are admitted while the staging
// budget allows (host buffers are copied into a device pool; device buffers
// are used in place and cost nothing).  Every admitted request joins the next
// batch; a batch is packed into one descriptor upload and one kernel launch
// per (op, degree, ...) group and issued on one of `slots` streams.  While a
// slot's batch runs, newly arrived requests are admitted into the next free
// slot, so work from later arrivals overlaps on the device.  The scheduler
// observes completion per slot (event query, polled every 20 us while a
// batch is in flight) and releases the batch's memory.  Released staging
// blocks stay cached for later requests; the pool's whole footprint (blocks
// in use and cached) plus the descriptor buffers is kept within the memory
// budget by freeing cached blocks before a new one is allocated.  Its thread runs with
// a 1 us timer slack: with Linux's default 50 us slack each 20 us poll slept
// ~70 us, which showed in every request's latency (p50 at half load 84 ->
// 68 us when the slack was cut, measured side by side).  Waiter threads
// blocking on each slot's event were measured too and were slower than the
// poll: the extra thread hand-off cost more than it saved.
#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "eltwise_host.hpp"
#include "ntt_host.hpp"

using axhel_dev::NttJobDev;
using namespace axhel_host;

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { ST_QUEUED, ST_RUNNING, ST_DONE };

struct Req {
    uint64_t id = 0;
    axhel_hip_request r{};
    std::vector<const axhel_hip_plan *> plans;
    std::vector<uint64_t> moduli, scalars;
    uint64_t words = 0;
    uint64_t bytes = 0;  // staging bytes the request needs (admission)
    uint64_t held = 0;   // bytes of the pool blocks it holds (a reused block may be larger)
    bool h_op1 = false, h_op2 = false, h_res = false;
    Clock::time_point arrival, done;
    int state = ST_QUEUED;
    axhel_hip_status status = AXHEL_HIP_OK;
    bool waited = false;
    // device-side views of the buffers
    const uint64_t *d_op1 = nullptr, *d_op2 = nullptr;
    uint64_t *d_res = nullptr;
    std::vector<std::pair<uint64_t, void *>> blocks;  // pool blocks (bytes, ptr)
};

struct Slot {
    bool busy = false;
    std::vector<Req *> batch;
    hipStream_t stream = nullptr;
    hipEvent_t event = nullptr;
    void *h_desc = nullptr;  // pinned descriptor staging
    void *d_desc = nullptr;
    size_t desc_cap = 0;
    Clock::time_point null_done;  // NULL backend
    hipError_t error = hipSuccess;
};

bool is_ntt(int op) { return op == AXHEL_HIP_OP_NTT_FORWARD || op == AXHEL_HIP_OP_NTT_INVERSE; }

// Latencies of completed requests in bounded memory: every sample up to cap
// (2^20 by default, 8 MiB), then a uniform sample of cap of them (reservoir
// sampling, Vitter's Algorithm R) from which the percentiles are read; count,
// sum and max stay exact.  A server that runs for days keeps its memory.
struct LatencyRecord {
    size_t cap = size_t(1) << 20;
    std::vector<double> sample;
    uint64_t count = 0;
    double sum = 0, max = 0;
    uint64_t rng = 0x9E3779B97F4A7C15ull;

    void add(double us) {
        ++count;
        sum += us;
        max = std::max(max, us);
        if (sample.size() < cap) {
            sample.push_back(us);
            return;
        }
        rng ^= rng << 13;  // xorshift64
        rng ^= rng >> 7;
        rng ^= rng << 17;
        const uint64_t j = rng % count;
        if (j < cap) sample[j] = us;
    }
    void clear() {
        sample.clear();
        count = 0;
        sum = max = 0;
    }
};

}  // namespace

struct axhel_hip_server {
    axhel_hip_serve_config cfg{};
    bool device = true;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    uint64_t batch_count = 0;  // batches issued (fault injection)
    uint64_t desc_bytes = 0;   // descriptor buffers of every slot, charged to the budget beside in_use
    std::deque<Req *> waiting;
    std::unordered_map<uint64_t, Req *> reqs;
    uint64_t next_id = 1;
    uint64_t budget = 0, in_use = 0, peak = 0;
    uint64_t alloc_count = 0;
    std::vector<Slot> slots;
    std::thread worker;
    bool stopping = false;
    int active_waiters = 0;  // threads inside axhel_hip_server_wait
    // pool: cached device blocks by size; every block with its size
    std::multimap<uint64_t, void *> free_blocks;
    std::unordered_map<void *, uint64_t> all_blocks;
    uint64_t pool_bytes = 0, pool_peak = 0;
    // plan cache
    std::map<std::tuple<uint64_t, uint64_t, uint64_t>, axhel_hip_plan *> plans;
    // stats
    axhel_hip_serve_stats st{};
    LatencyRecord lat;
    bool have_first = false;
    Clock::time_point first_arrival, last_done;
};

namespace {

// ------------------------------------------------------------ staging pool

void pool_free_block(axhel_hip_server *s, void *b) {
    if (s->device) (void)hipFree(b);
    else std::free(b);
}

enum class Alloc { OK, OVER_BUDGET, FAILED };

// A block of at least `bytes`: a cached one up to twice that size, else a
// new one.  Before allocating, cached blocks are freed (largest first) until
// the pool plus the descriptor buffers fit the budget with the new block;
// none of them is read by queued work (a block returns to the cache when its
// batch has completed).  If they still do not fit, OVER_BUDGET, unless
// may_exceed.  hipFree waits for the device, so trimming stalls the
// scheduler for the batches in flight; it happens only when the size mix
// changes.  Called with the server mutex held.
Alloc pool_alloc(axhel_hip_server *s, uint64_t bytes, bool may_exceed, void **p, uint64_t *got) {
    ++s->alloc_count;
    if (s->cfg.fail_every_alloc && s->alloc_count % s->cfg.fail_every_alloc == 0) return Alloc::FAILED;
    auto it = s->free_blocks.lower_bound(bytes);
    if (it != s->free_blocks.end() && it->first <= bytes * 2) {
        *p = it->second;
        *got = it->first;
        s->free_blocks.erase(it);
        return Alloc::OK;
    }
    while (!s->free_blocks.empty() && s->pool_bytes + s->desc_bytes + bytes > s->budget) {
        auto big = std::prev(s->free_blocks.end());
        pool_free_block(s, big->second);
        s->all_blocks.erase(big->second);
        s->pool_bytes -= big->first;
        s->free_blocks.erase(big);
    }
    if (!may_exceed && s->pool_bytes + s->desc_bytes + bytes > s->budget) return Alloc::OVER_BUDGET;
    void *d = nullptr;
    if (s->device) {
        if (hipMalloc(&d, bytes) != hipSuccess) return Alloc::FAILED;
    } else {
        d = std::malloc(bytes);
        if (!d) return Alloc::FAILED;
    }
    s->all_blocks.emplace(d, bytes);
    s->pool_bytes += bytes;
    s->pool_peak = std::max(s->pool_peak, s->pool_bytes);
    *p = d;
    *got = bytes;
    return Alloc::OK;
}

void pool_release(axhel_hip_server *s, Req *R) {
    for (auto &b : R->blocks) s->free_blocks.emplace(b.first, b.second);
    R->blocks.clear();
    s->in_use -= R->held;
    R->held = 0;
}

// Reserves staging for R's host buffers.  On anything but OK the blocks
// taken so far go back to the cache.
Alloc stage_alloc(axhel_hip_server *s, Req *R, bool may_exceed) {
    const uint64_t wb = R->words * sizeof(uint64_t);
    Alloc a = Alloc::OK;
    auto get = [&](void **p) -> bool {
        // block size rounded up to 1 MiB so blocks are reused across sizes
        const uint64_t sz = (wb + (1u << 20) - 1) & ~(uint64_t)((1u << 20) - 1);
        uint64_t got = 0;
        if ((a = pool_alloc(s, sz, may_exceed, p, &got)) != Alloc::OK) return false;
        R->blocks.emplace_back(got, *p);
        R->held += got;
        return true;
    };
    void *p = nullptr;
    R->d_op1 = R->r.operand1;
    R->d_op2 = R->r.operand2;
    R->d_res = R->r.result;
    bool ok = true;
    if (R->h_op1 && (ok = get(&p))) R->d_op1 = (const uint64_t *)p;
    if (ok && R->h_op2 && (ok = get(&p))) R->d_op2 = (const uint64_t *)p;
    if (ok && R->h_res) {
        if (R->r.result == R->r.operand1 && R->h_op1) {
            R->d_res = const_cast<uint64_t *>(R->d_op1);
        } else if (R->r.result == R->r.operand2 && R->h_op2) {
            R->d_res = const_cast<uint64_t *>(R->d_op2);
        } else if ((ok = get(&p))) {
            R->d_res = (uint64_t *)p;
        }
    }
    if (!ok) {
        for (auto &b : R->blocks) s->free_blocks.emplace(b.first, b.second);
        R->blocks.clear();
        R->held = 0;
    }
    return a;
}

// ------------------------------------------------------------- device batch

// The slot's descriptor buffers, at least `bytes` long.  Called with the
// server mutex released; the slot is idle (its previous batch completed), so
// its old buffers are no longer read by the device.  The device buffer's
// size is charged to the budget.
hipError_t ensure_desc(axhel_hip_server *s, Slot &sl, size_t bytes) {
    if (sl.desc_cap >= bytes) return hipSuccess;
    size_t cap = std::max<size_t>(bytes, 64 << 10);
    if (sl.h_desc) (void)hipHostFree(sl.h_desc);
    if (sl.d_desc) (void)hipFree(sl.d_desc);
    sl.h_desc = sl.d_desc = nullptr;
    const size_t old_cap = sl.desc_cap;
    sl.desc_cap = 0;
    hipError_t e = hipHostMalloc(&sl.h_desc, cap, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&sl.d_desc, cap);
    if (e == hipSuccess) sl.desc_cap = cap;
    std::lock_guard<std::mutex> lk(s->mu);
    s->desc_bytes = s->desc_bytes - old_cap + sl.desc_cap;
    s->peak = std::max(s->peak, s->in_use + s->desc_bytes);
    return e;
}

// Issues a whole batch on the slot's stream: staging copies in, one
// descriptor upload, one launch per group, copies out, completion event.
hipError_t launch_batch(axhel_hip_server *s, Slot &sl) {
    const hipStream_t st = sl.stream;
    hipError_t e = hipSuccess;
    for (Req *R : sl.batch) {
        const uint64_t wb = R->words * sizeof(uint64_t);
        if (R->h_op1 && (e = hipMemcpyAsync((void *)R->d_op1, R->r.operand1, wb, hipMemcpyHostToDevice, st)))
            return e;
        if (R->h_op2 && (e = hipMemcpyAsync((void *)R->d_op2, R->r.operand2, wb, hipMemcpyHostToDevice, st)))
            return e;
    }
    // group keys -> limb descriptors
    using NttKey = std::tuple<uint32_t, int, int, int>;  // logn, inverse, normalize, vec16
    using EltKey = std::tuple<int, int, uint64_t>;       // op, factor, degree
    std::map<NttKey, std::vector<NttJobDev>> ntt;
    std::map<EltKey, std::vector<EltLimbDev>> elt;
    for (Req *R : sl.batch) {
        const uint64_t n = R->r.degree;
        const int op = R->r.op;
        if (is_ntt(op)) {
            const bool inv = op == AXHEL_HIP_OP_NTT_INVERSE;
            const bool norm = inv ? R->r.output_mod_factor != 2 : R->r.output_mod_factor != 4;
            const bool v16 = aligned16(R->d_op1) && aligned16(R->d_res) && (n % 2 == 0);
            auto &v = ntt[NttKey(R->plans[0]->logn, inv ? 1 : 0, norm ? 1 : 0, v16 ? 1 : 0)];
            for (uint32_t l = 0; l < R->r.limbs; ++l)
                v.push_back(make_job(R->plans[l], R->d_op1 + l * n, R->d_res + l * n, inv));
        } else {
            int eop = 0, f = 1;
            switch (op) {
                case AXHEL_HIP_OP_ADD: eop = ELT_ADD_VV; break;
                case AXHEL_HIP_OP_SUB: eop = ELT_SUB_VV; break;
                case AXHEL_HIP_OP_MUL: eop = ELT_MUL; f = (int)R->r.input_mod_factor; break;
                case AXHEL_HIP_OP_FMA:
                    eop = R->r.operand2 ? ELT_FMA : ELT_FMA_NULL;
                    f = (int)R->r.input_mod_factor;
                    break;
                default: {  // reduce: branch order of eltwise-reduce-mod.cpp:32-128 (identity checked at submit)
                    const uint64_t fi = R->r.input_mod_factor, fo = R->r.output_mod_factor;
                    if (fi == 2 && fo == 1) eop = ELT_RED_2_1;
                    else if (fi == 4 && fo == 1) eop = ELT_RED_4_1;
                    else if (fi == 4 && fo == 2) eop = ELT_RED_4_2;
                    else eop = ELT_RED_BARRETT;
                }
            }
            auto &v = elt[EltKey(eop, f, n)];
            for (uint32_t l = 0; l < R->r.limbs; ++l) {
                EltLimbDev d;
                const uint64_t q = R->moduli[l];
                d.r = R->d_res + l * n;
                d.a = R->d_op1 + l * n;
                d.b = R->d_op2 ? R->d_op2 + l * n : nullptr;
                uint64_t sc = R->scalars.empty() ? 0 : R->scalars[l];
                if (op == AXHEL_HIP_OP_FMA) {  // ReduceInputNative<F>(op2), eltwise-fma-mod.cpp:32
                    if (f == 8 && sc >= (q << 2)) sc -= q << 2;
                    if (f >= 4 && sc >= (q << 1)) sc -= q << 1;
                    if (f >= 2 && sc >= q) sc -= q;
                }
                d.s = sc;
                d.bc = barrett_for(q);
                v.push_back(d);
            }
        }
    }
    size_t bytes = 0;
    for (auto &kv : ntt) bytes += kv.second.size() * sizeof(NttJobDev);
    for (auto &kv : elt) bytes += kv.second.size() * sizeof(EltLimbDev);
    bytes += 64 * (ntt.size() + elt.size());
    if ((e = ensure_desc(s, sl, bytes)) != hipSuccess) return e;
    if (s->cfg.fail_every_batch && (++s->batch_count % s->cfg.fail_every_batch) == 0)
        return hipErrorOutOfMemory;  // injected, after the staging copies were queued
    char *h = (char *)sl.h_desc;
    size_t off = 0;
    std::vector<std::pair<const void *, size_t>> places;
    for (auto &kv : ntt) {
        std::memcpy(h + off, kv.second.data(), kv.second.size() * sizeof(NttJobDev));
        places.emplace_back((char *)sl.d_desc + off, kv.second.size());
        off += (kv.second.size() * sizeof(NttJobDev) + 63) & ~(size_t)63;
    }
    for (auto &kv : elt) {
        std::memcpy(h + off, kv.second.data(), kv.second.size() * sizeof(EltLimbDev));
        places.emplace_back((char *)sl.d_desc + off, kv.second.size());
        off += (kv.second.size() * sizeof(EltLimbDev) + 63) & ~(size_t)63;
    }
    if (off && (e = hipMemcpyAsync(sl.d_desc, sl.h_desc, off, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    static const NttJobDev none{};
    size_t pi = 0;
    for (auto &kv : ntt) {
        const uint32_t logn = std::get<0>(kv.first);
        const NttConfig &cfg = config_for_launch(logn, kv.second.size());
        if ((e = prepare_config(cfg)) != hipSuccess) return e;
        bool relax_ok = true;
        for (const NttJobDev &j : kv.second) relax_ok = relax_ok && relax_ok_modulus(j.q);
        e = launch_ntt(cfg, logn, std::get<1>(kv.first) != 0, std::get<2>(kv.first) != 0,
                       (const NttJobDev *)places[pi].first, none, (uint32_t)places[pi].second, st, -1,
                       std::get<3>(kv.first) != 0, nullptr, relax_ok);
        if (e != hipSuccess) return e;
        ++pi;
    }
    for (auto &kv : elt) {
        e = launch_elt_batch(std::get<0>(kv.first), std::get<1>(kv.first), (const EltLimbDev *)places[pi].first,
                             (uint32_t)places[pi].second, std::get<2>(kv.first), st);
        if (e != hipSuccess) return e;
        ++pi;
    }
    for (Req *R : sl.batch) {
        if (R->h_res &&
            (e = hipMemcpyAsync(R->r.result, R->d_res, R->words * sizeof(uint64_t), hipMemcpyDeviceToHost, st)))
            return e;
    }
    return hipEventRecord(sl.event, st);
}

void finish_req(axhel_hip_server *s, Req *R, axhel_hip_status status) {
    R->status = status;
    R->state = ST_DONE;
    R->done = Clock::now();
    switch (status) {
        case AXHEL_HIP_OK: {
            ++s->st.completed;
            s->st.limbs_completed += R->r.limbs;
            s->lat.add(std::chrono::duration<double, std::micro>(R->done - R->arrival).count());
            s->last_done = R->done;
            break;
        }
        case AXHEL_HIP_TIMEOUT: ++s->st.timed_out; break;
        case AXHEL_HIP_CANCELLED: ++s->st.cancelled; break;
        case AXHEL_HIP_REJECTED: ++s->st.rejected; break;
        default: ++s->st.failed; break;
    }
}

bool slot_done(axhel_hip_server *s, Slot &sl) {
    if (!s->device) return sl.error != hipSuccess || Clock::now() >= sl.null_done;
    if (sl.error != hipSuccess) return true;
    const hipError_t q = hipEventQuery(sl.event);
    if (q == hipErrorNotReady) return false;
    if (q != hipSuccess) sl.error = q;
    return true;
}

void scheduler(axhel_hip_server *s) {
    prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);  // 1 us: the 20 us poll below sleeps 20 us, not ~70
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progressed = false;
        bool any_busy = false;
        // 1. completions
        for (auto &sl : s->slots) {
            if (!sl.busy) continue;
            if (slot_done(s, sl)) {
                const axhel_hip_status stt = sl.error == hipSuccess             ? AXHEL_HIP_OK
                                             : sl.error == hipErrorOutOfMemory ? AXHEL_HIP_NO_MEMORY
                                                                                : AXHEL_HIP_HIP_ERROR;
                for (Req *R : sl.batch) {
                    pool_release(s, R);
                    finish_req(s, R, stt);
                }
                sl.batch.clear();
                sl.busy = false;
                sl.error = hipSuccess;
                progressed = true;
                s->cv_done.notify_all();
            } else {
                any_busy = true;
            }
        }
        // 2. queued requests past their deadline
        const auto now = Clock::now();
        for (auto it = s->waiting.begin(); it != s->waiting.end();) {
            Req *R = *it;
            if (R->r.timeout_us > 0 && now - R->arrival > std::chrono::microseconds(R->r.timeout_us)) {
                finish_req(s, R, AXHEL_HIP_TIMEOUT);
                it = s->waiting.erase(it);
                progressed = true;
                s->cv_done.notify_all();
            } else {
                ++it;
            }
        }
        // 3. admission into a free slot
        Slot *free_slot = nullptr;
        for (auto &sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (free_slot && !s->waiting.empty()) {
            uint64_t limbs = 0;
            while (!s->waiting.empty()) {
                Req *R = s->waiting.front();
                // head of line waits for memory (staging plus the descriptor
                // buffers); with no staging in use the request, no larger than
                // the budget (checked at submit), goes ahead
                if (s->in_use + s->desc_bytes + R->bytes > s->budget && s->in_use > 0) break;
                if (!free_slot->batch.empty() && limbs + R->r.limbs > s->cfg.max_batch_limbs) break;
                s->waiting.pop_front();
                // the pool may pass the budget only for a request that runs alone
                const Alloc a = stage_alloc(s, R, s->in_use == 0);
                if (a == Alloc::OVER_BUDGET) {  // waits for running batches to return their blocks
                    s->waiting.push_front(R);
                    break;
                }
                if (a == Alloc::FAILED) {
                    finish_req(s, R, AXHEL_HIP_NO_MEMORY);
                    s->cv_done.notify_all();
                    continue;
                }
                s->in_use += R->held;
                s->peak = std::max(s->peak, s->in_use + s->desc_bytes);
                R->state = ST_RUNNING;
                free_slot->batch.push_back(R);
                limbs += R->r.limbs;
            }
            if (!free_slot->batch.empty()) {
                free_slot->busy = true;
                ++s->st.batches;
                progressed = true;
                any_busy = true;
                if (s->device) {
                    // the slot's requests are owned by the scheduler now; launch unlocked
                    lk.unlock();
                    const hipError_t e = launch_batch(s, *free_slot);
                    if (e != hipSuccess) {
                        // copies or kernels queued before the failure may still
                        // read or write the batch's staging blocks: drain the
                        // stream before they return to the pool
                        (void)hipStreamSynchronize(free_slot->stream);
                    }
                    lk.lock();
                    if (e != hipSuccess) {
                        set_last_error("serve batch", e);
                        free_slot->error = e;
                    }
                } else {
                    if (s->cfg.fail_every_batch && (++s->batch_count % s->cfg.fail_every_batch) == 0)
                        free_slot->error = hipErrorOutOfMemory;  // injected
                    free_slot->null_done = Clock::now() + std::chrono::microseconds(s->cfg.null_latency_us);
                }
            }
        }
        if (s->stopping && s->waiting.empty() && !any_busy) break;
        if (!progressed) {
            if (any_busy)
                s->cv_work.wait_for(lk, std::chrono::microseconds(20));
            else if (!s->waiting.empty())
                s->cv_work.wait_for(lk, std::chrono::microseconds(200));  // memory-blocked or deadlines
            else
                s->cv_work.wait(lk);
        }
    }
}

axhel_hip_status validate(const axhel_hip_request *r) {
    if (!r || r->limbs == 0 || r->degree < 2 || (r->degree & (r->degree - 1)) || r->degree > (1ull << kMaxLogN))
        return AXHEL_HIP_INVALID_ARGUMENT;
    if (!r->result || !r->operand1) return AXHEL_HIP_INVALID_ARGUMENT;
    switch (r->op) {
        case AXHEL_HIP_OP_NTT_FORWARD:
        case AXHEL_HIP_OP_NTT_INVERSE:
            if (!r->plans) return AXHEL_HIP_INVALID_ARGUMENT;
            for (uint32_t l = 0; l < r->limbs; ++l)
                if (!r->plans[l] || r->plans[l]->n != r->degree) return AXHEL_HIP_INVALID_ARGUMENT;
            return AXHEL_HIP_OK;
        case AXHEL_HIP_OP_ADD:
        case AXHEL_HIP_OP_SUB:
        case AXHEL_HIP_OP_MUL:
        case AXHEL_HIP_OP_FMA:
        case AXHEL_HIP_OP_REDUCE:
            if (!r->moduli) return AXHEL_HIP_INVALID_ARGUMENT;
            for (uint32_t l = 0; l < r->limbs; ++l)
                if (r->moduli[l] < 2) return AXHEL_HIP_INVALID_ARGUMENT;
            if ((r->op == AXHEL_HIP_OP_ADD || r->op == AXHEL_HIP_OP_SUB || r->op == AXHEL_HIP_OP_MUL) && !r->operand2)
                return AXHEL_HIP_INVALID_ARGUMENT;
            if (r->op == AXHEL_HIP_OP_MUL && r->input_mod_factor != 1 && r->input_mod_factor != 2 &&
                r->input_mod_factor != 4)
                return AXHEL_HIP_UNSUPPORTED_FACTOR;
            if (r->op == AXHEL_HIP_OP_FMA) {
                if (!r->scalars) return AXHEL_HIP_INVALID_ARGUMENT;
                const uint64_t f = r->input_mod_factor;
                if (f != 1 && f != 2 && f != 4 && f != 8) return AXHEL_HIP_UNSUPPORTED_FACTOR;
            }
            if (r->op == AXHEL_HIP_OP_REDUCE) {
                // The arithmetic branches of eltwise-reduce-mod.cpp:32-128.
                // input_mod_factor 0 selects each limb's own Barrett branch
                // (the reference's in_mod_factor == modulus, which one factor
                // cannot express for limbs with different moduli).  The
                // identity branches are plain copies and are not served.
                const uint64_t fi = r->input_mod_factor, fo = r->output_mod_factor;
                const bool arith = (fi == 2 && fo == 1) || (fi == 4 && fo == 1) || (fi == 4 && fo == 2);
                bool barrett = fi == 0;
                if (!barrett && r->limbs == 1 && fi == r->moduli[0] && fi > 4) barrett = true;
                if (!arith && !barrett) return AXHEL_HIP_UNSUPPORTED_FACTOR;
            }
            return AXHEL_HIP_OK;
    }
    return AXHEL_HIP_INVALID_ARGUMENT;
}

// Linear interpolation between the closest ranks of a sorted sample.
double percentile_sorted(const std::vector<double> &v, double p) {
    if (v.empty()) return 0.0;
    const double pos = p * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

axhel_hip_status axhel_hip_server_create(const axhel_hip_serve_config *config, axhel_hip_server **server) {
    if (!server) return AXHEL_HIP_INVALID_ARGUMENT;
    *server = nullptr;
    axhel_hip_server *s = new (std::nothrow) axhel_hip_server;
    if (!s) return AXHEL_HIP_NO_MEMORY;
    if (config) s->cfg = *config;
    if (s->cfg.max_batch_limbs == 0) s->cfg.max_batch_limbs = 8192;
    if (s->cfg.slots == 0) s->cfg.slots = 4;  // one stream per hardware queue of a process (HIP's default 4)
    if (s->cfg.latency_samples) s->lat.cap = (size_t)std::min<uint64_t>(s->cfg.latency_samples, 1ull << 28);
    s->device = s->cfg.backend == AXHEL_HIP_BACKEND_DEVICE;
    if (s->device && axhel_hip_device_count() == 0) {
        delete s;
        return AXHEL_HIP_NO_DEVICE;
    }
    s->budget = s->cfg.memory_budget_bytes;
    if (s->budget == 0) {
        size_t fr = 0, tot = 0;
        if (s->device && hipMemGetInfo(&fr, &tot) == hipSuccess) s->budget = fr / 2;
        else s->budget = 1ull << 30;
    }
    s->slots.resize(s->cfg.slots);
    if (s->device) {
        for (auto &sl : s->slots) {
            hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.event, hipEventDisableTiming);
            if (e != hipSuccess) {
                set_last_error("server create", e);
                for (auto &x : s->slots) {
                    if (x.event) (void)hipEventDestroy(x.event);
                    if (x.stream) (void)hipStreamDestroy(x.stream);
                }
                delete s;
                return AXHEL_HIP_HIP_ERROR;
            }
        }
    }
    s->worker = std::thread(scheduler, s);
    *server = s;
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_server_destroy(axhel_hip_server *s) {
    if (!s) return AXHEL_HIP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        for (Req *R : s->waiting) finish_req(s, R, AXHEL_HIP_CANCELLED);
        s->waiting.clear();
        s->cv_done.notify_all();
    }
    s->cv_work.notify_all();
    s->worker.join();
    {
        // let threads still inside wait() observe their final status and leave
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [&] { return s->active_waiters == 0; });
    }
    for (auto &kv : s->reqs) delete kv.second;
    for (auto &sl : s->slots) {
        if (sl.event) (void)hipEventDestroy(sl.event);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_desc) (void)hipHostFree(sl.h_desc);
        if (sl.d_desc) (void)hipFree(sl.d_desc);
    }
    for (auto &kv : s->all_blocks) pool_free_block(s, kv.first);
    for (auto &kv : s->plans) axhel_hip_ntt_plan_destroy(kv.second);
    delete s;
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_server_plan(axhel_hip_server *s, uint64_t degree, uint64_t modulus,
                                       uint64_t root_of_unity, const axhel_hip_plan **plan) {
    if (!s || !plan) return AXHEL_HIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto key = std::make_tuple(degree, modulus, root_of_unity);
    auto it = s->plans.find(key);
    if (it != s->plans.end()) {
        *plan = it->second;
        return AXHEL_HIP_OK;
    }
    axhel_hip_plan *p = nullptr;
    axhel_hip_status st = axhel_hip_ntt_plan_create(degree, modulus, root_of_unity, &p);
    if (st != AXHEL_HIP_OK) return st;
    s->plans[key] = p;
    *plan = p;
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_server_submit(axhel_hip_server *s, const axhel_hip_request *request, uint64_t *ticket) {
    if (!s || !ticket) return AXHEL_HIP_INVALID_ARGUMENT;
    const axhel_hip_status v = validate(request);
    if (v != AXHEL_HIP_OK) return v;
    Req *R = new (std::nothrow) Req;
    if (!R) return AXHEL_HIP_NO_MEMORY;
    R->r = *request;
    R->words = (uint64_t)request->limbs * request->degree;
    if (is_ntt(request->op)) R->plans.assign(request->plans, request->plans + request->limbs);
    else R->moduli.assign(request->moduli, request->moduli + request->limbs);
    if (request->op == AXHEL_HIP_OP_FMA) R->scalars.assign(request->scalars, request->scalars + request->limbs);
    if (s->device) {
        R->h_op1 = !is_device_pointer(request->operand1);
        R->h_op2 = request->operand2 && !is_device_pointer(request->operand2);
        R->h_res = !is_device_pointer(request->result);
    } else {
        R->h_op1 = R->h_res = true;  // the NULL backend stages everything: exercises admission
        R->h_op2 = request->operand2 != nullptr;
    }
    uint64_t nbuf = (R->h_op1 ? 1 : 0) + (R->h_op2 ? 1 : 0);
    if (R->h_res && !((request->result == request->operand1 && R->h_op1) ||
                      (request->result == request->operand2 && R->h_op2)))
        ++nbuf;
    const uint64_t block = ((R->words * 8) + (1u << 20) - 1) & ~(uint64_t)((1u << 20) - 1);
    R->bytes = nbuf * block;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
        delete R;
        return AXHEL_HIP_REJECTED;
    }
    if (R->bytes > s->budget) {
        ++s->st.rejected;
        delete R;
        return AXHEL_HIP_REJECTED;
    }
    R->id = s->next_id++;
    R->arrival = Clock::now();
    if (!s->have_first) {
        s->have_first = true;
        s->first_arrival = R->arrival;
    }
    ++s->st.submitted;
    s->reqs[R->id] = R;
    s->waiting.push_back(R);
    *ticket = R->id;
    s->cv_work.notify_one();
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_server_wait(axhel_hip_server *s, uint64_t ticket, int64_t timeout_us) {
    if (!s) return AXHEL_HIP_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(ticket) == s->reqs.end()) return AXHEL_HIP_INVALID_ARGUMENT;
    // Looked up by ticket on every wake-up: another thread waiting on the same
    // ticket may have collected (and freed) the request meanwhile.
    auto pred = [&] {
        auto f = s->reqs.find(ticket);
        return f == s->reqs.end() || f->second->state == ST_DONE;
    };
    ++s->active_waiters;
    bool done = true;
    if (timeout_us < 0) s->cv_done.wait(lk, pred);
    else done = s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), pred);
    --s->active_waiters;
    s->cv_done.notify_all();
    if (!done) return AXHEL_HIP_PENDING;
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return AXHEL_HIP_INVALID_ARGUMENT;  // collected by another waiter
    Req *R = it->second;
    const axhel_hip_status st = R->status;
    s->reqs.erase(it);
    delete R;
    return st;
}

axhel_hip_status axhel_hip_server_cancel(axhel_hip_server *s, uint64_t ticket) {
    if (!s) return AXHEL_HIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return AXHEL_HIP_INVALID_ARGUMENT;
    Req *R = it->second;
    if (R->state != ST_QUEUED) return AXHEL_HIP_REJECTED;
    s->waiting.erase(std::find(s->waiting.begin(), s->waiting.end(), R));
    finish_req(s, R, AXHEL_HIP_CANCELLED);
    s->cv_done.notify_all();
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_server_stats(axhel_hip_server *s, axhel_hip_serve_stats *out) {
    if (!s || !out) return AXHEL_HIP_INVALID_ARGUMENT;
    std::vector<double> sample;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->st;
        out->bytes_in_use = s->in_use;
        out->peak_bytes_in_use = s->peak;
        out->descriptor_bytes = s->desc_bytes;
        out->pool_bytes = s->pool_bytes;
        out->peak_pool_bytes = s->pool_peak;
        out->mean_us = s->lat.count ? s->lat.sum / (double)s->lat.count : 0;
        out->max_us = s->lat.max;
        out->first_arrival_to_last_completion_s =
            s->have_first && s->st.completed ? std::chrono::duration<double>(s->last_done - s->first_arrival).count()
                                             : 0.0;
        sample = s->lat.sample;
    }
    // sorted outside the lock: the scheduler does not wait for it
    std::sort(sample.begin(), sample.end());
    out->p50_us = percentile_sorted(sample, 0.50);
    out->p99_us = percentile_sorted(sample, 0.99);
    return AXHEL_HIP_OK;
}

axhel_hip_status axhel_hip_server_reset_stats(axhel_hip_server *s) {
    if (!s) return AXHEL_HIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    s->st = axhel_hip_serve_stats{};
    s->lat.clear();
    s->have_first = false;
    s->peak = s->in_use + s->desc_bytes;
    s->pool_peak = s->pool_bytes;
    return AXHEL_HIP_OK;
}

}  // extern "C"
