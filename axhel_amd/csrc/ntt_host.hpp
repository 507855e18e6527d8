// Host-side NTT internals shared by the plain, batched and serving paths.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "common.hpp"
#include "ntt_kernels.hpp"

// Plan: one (degree, modulus, root) with resident device tables.
struct axhel_hip_plan {
    uint64_t n = 0;
    uint32_t logn = 0;
    uint64_t q = 0;
    int device = 0;
    std::vector<uint64_t> h_fwd, h_inv;  // AoS {operand, quotient}, 2n words each
    // External tables: the caller's arrays, borrowed for the plan's life as
    // the reference borrows them (ntt.hpp:34-37).  h_fwd/h_inv hold what the
    // device tables were last loaded from; the per-call transforms reload a
    // table that the caller has changed since.  ext_mu serializes the check,
    // the reload and the queuing of the transform's kernels.
    const uint64_t *ext_fwd = nullptr, *ext_inv = nullptr;
    mutable std::mutex ext_mu;
    uint64_t ninv = 0, ninv_q = 0;
    uint64_t sroot = 0, sroot_q = 0;  // CS(Shoup(inv[n-1], n^-1), q) and its quotient
    ulonglong2 *d_fwd = nullptr;
    ulonglong2 *d_inv = nullptr;
};

namespace axhel_host {

constexpr uint32_t kMaxLogN = 20;

// One compiled pass kernel.
struct PassKernel {
    const void *fn = nullptr;
    // forward kernels: the same geometry on the relaxed butterfly schedule
    // (normalized outputs only, every modulus below 2^60); null otherwise
    const void *fn_relaxed = nullptr;
    int log_g = 0;
    int log_e = 0;
    int gpw = 0;
    bool col = false;
    bool multi = false;
    bool stack = false;  // ROW pass over stacks of same-plan jobs
    int threads = 0;
    size_t lds_bytes = 0;          // fn's dynamic LDS
    size_t lds_bytes_relaxed = 0;  // fn_relaxed's (its tile layout may differ)
};

// One transform configuration for a degree: the passes in execution order,
// and the fused polynomial-product kernel on the same ROW geometry (fn null
// when the configuration has none).
struct NttConfig {
    axhel_hip_ntt_config pub{};
    int npass = 0;
    PassKernel fwd[2];
    PassKernel inv[2];
    PassKernel fused;
    // Two-pass configurations: the ROW pass over stacks of jobs sharing one
    // plan (forward pass 1, inverse pass 0), used when a stack table is given.
    PassKernel fwd_stack, inv_stack;
    // Fused polynomial-product kernel over the same stacks (fn null: none).
    PassKernel fused_stack;
};

// Jobs of one launch regrouped for the stacked ROW pass: `count` stacks of
// `gpw` slots, each stack's jobs sharing a plan, empty slots with dst null;
// the jobs left out of every stack (`rest`, rest_count of them) run the
// per-limb ROW kernel beside the stacked one.
struct StackTable {
    const axhel_dev::NttJobDev *jobs = nullptr;
    uint32_t count = 0;
    int gpw = 0;
    const axhel_dev::NttJobDev *rest = nullptr;
    uint32_t rest_count = 0;
};

// Builds the stacked job table from `jobs` ordered so that jobs of one plan
// are adjacent (runs of equal table and modulus): each plan's jobs fill
// whole stacks, and its last partial stack is padded with empty slots when
// at most one slot in eight stays empty, else its jobs go to `rest`.
// Returns false, leaving both outputs empty, when no stack forms.
bool build_stacks(const axhel_dev::NttJobDev *jobs, uint32_t count, int gpw, std::vector<axhel_dev::NttJobDev> &out,
                  uint32_t &stacks, std::vector<axhel_dev::NttJobDev> &rest);

// Compiled configurations for log2(degree) (empty for unsupported degrees).
const std::vector<NttConfig> &configs_for(uint32_t logn);
// Configuration currently selected for log2(degree) (batches, every multi-limb path).
const NttConfig &config_for(uint32_t logn);
// Configuration of single-limb calls: the tuning table's latency choice, or
// config_for's once a configuration has been selected explicitly.
const NttConfig &config_for_single(uint32_t logn);
// Configuration of launches of 2 limbs up to kLatencyWords words: the
// tuning table's small-launch entry, else config_for_single's.
const NttConfig &config_for_small(uint32_t logn);
// Configuration of one launch over `count` limbs.  One limb runs the
// single-call configuration; up to kLatencyWords words a launch fills too
// little of the GPU to be throughput-bound and runs the small-launch
// configuration (latency shapes, measured faster up to 32 limbs at 2^16 and
// 64+ at 2^13 and 2^14); larger launches run the batch configuration.
constexpr uint64_t kLatencyWords = 1ull << 21;
const NttConfig &config_for_launch(uint32_t logn, uint64_t count);
// Group count per stack of a configuration's stacked ROW pass (0: none; a
// one-row stack would share its twiddles with nothing).
inline int stack_gpw_of(const NttConfig &c) { return c.fwd_stack.fn && c.fwd_stack.gpw > 1 ? c.fwd_stack.gpw : 0; }

// Raises per-kernel limits (dynamic LDS above 64 KiB) for every pass of a
// configuration; done before graph capture so capture records launches only.
hipError_t prepare_config(const NttConfig &cfg);

// Device job descriptor for plan + pointers.
axhel_dev::NttJobDev make_job(const axhel_hip_plan *p, const uint64_t *src, uint64_t *dst, bool inverse);

// Launches all passes of one transform family: `count` limbs of the same
// degree and direction whose descriptors are `jobs` (device table) or, when
// jobs == nullptr and count == 1, `inline_job`.  only_pass >= 0 launches
// that pass alone (per-kernel timing).  vec16: every src/dst is 16-byte
// aligned (enables 16-byte global accesses).  stacks: the same jobs as a
// stack table; the ROW pass of a two-pass configuration then runs stacked
// when the table was built for that pass's group count.  relax_ok: every
// job's modulus is below 2^60 (relax_ok_modulus), so a normalized forward may
// run on the relaxed butterfly schedule (same outputs, fewer subtractions).
hipError_t launch_ntt(const NttConfig &cfg, uint32_t logn, bool inverse, bool normalize,
                      const axhel_dev::NttJobDev *jobs, const axhel_dev::NttJobDev &inline_job, uint32_t count,
                      hipStream_t stream, int only_pass = -1, bool vec16 = false,
                      const StackTable *stacks = nullptr, bool relax_ok = false);

// Moduli for which the relaxed forward schedule keeps every word below 2^64
// (its words stay below 16q).
inline bool relax_ok_modulus(uint64_t q) { return q < (1ull << 60); }

// Whether a forward transform under cfg runs the relaxed schedule: its output
// is normalized, every modulus passes relax_ok_modulus (relax_ok), and every
// forward pass kernel of cfg has a relaxed twin.  launch_ntt decides by this,
// and so must any caller that runs part of a forward transform itself.
bool forward_relaxed(const NttConfig &cfg, bool normalize, bool relax_ok);

// Group count per stack of the stacked ROW pass for log2(degree) under the
// selected configuration (0: no stacked pass).
int stack_gpw(uint32_t logn);

}  // namespace axhel_host
