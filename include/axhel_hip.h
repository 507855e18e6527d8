/*
 * axhel_hip.h -- C ABI of the MI355X (gfx950) AXHEL library.
 *
 * Plain C: pointers, sizes and status codes only.  Streams are passed as
 * `void *` holding a hipStream_t (NULL = the default stream).
 *
 * Pointer rule for every compute entry point: operands and results may be
 * device pointers (hipMalloc / managed memory), which run asynchronously on
 * `stream`, or ordinary host pointers, which the library stages through
 * device memory and completes before returning.  Results may alias operands
 * exactly, as in the reference (tests/public/test_eltwise.cpp:439-516).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * massidonati/AXHEL).  Arithmetic is bit-identical to the reference's Native
 * backend, lazy outputs included.
 */
#ifndef AXHEL_HIP_H
#define AXHEL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum axhel_hip_status {
    AXHEL_HIP_OK = 0,
    /* Factor combination the reference does nothing for (silent no-op there:
       eltwise-mul-mod.cpp:83-93, eltwise-fma-mod.cpp:98-111,
       eltwise-reduce-mod.cpp:32-128).  Output is left untouched. */
    AXHEL_HIP_UNSUPPORTED_FACTOR = 1,
    AXHEL_HIP_INVALID_ARGUMENT = 2,
    AXHEL_HIP_NO_MEMORY = 3,
    AXHEL_HIP_HIP_ERROR = 4,
    AXHEL_HIP_NO_DEVICE = 5,
    AXHEL_HIP_REJECTED = 6,  /* serving: admission refused (memory budget / shutdown) */
    AXHEL_HIP_TIMEOUT = 7,   /* serving: a request dropped because it was not started by its deadline */
    AXHEL_HIP_CANCELLED = 8,
    AXHEL_HIP_PENDING = 9    /* serving: axhel_hip_server_wait gave up first; the request runs on */
} axhel_hip_status;

/* Human-readable text of a status, and the last HIP error string seen by the
   calling thread (empty if none). */
const char *axhel_hip_status_string(axhel_hip_status s);
const char *axhel_hip_last_error(void);

/* ------------------------------------------------------------------ device */

/* Number of visible GPUs; 0 when no GPU is present (no error). */
int axhel_hip_device_count(void);
/* 1 if p is device-accessible memory (hipMalloc/managed), else 0. */
int axhel_hip_is_device_pointer(const void *p);
axhel_hip_status axhel_hip_stream_synchronize(void *stream);

/* ------------------------------------------------------------- element-wise */

/* EltwiseAddMod vector-vector: include/axhel/eltwise/eltwise-add-mod.hpp:13
   (native body eltwise-add-mod.cpp:15-30). */
axhel_hip_status axhel_hip_add_mod_vv(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                                      uint64_t n, uint64_t modulus, void *stream);
/* EltwiseAddMod vector-scalar: eltwise-add-mod.hpp:17 (eltwise-add-mod.cpp:33-47). */
axhel_hip_status axhel_hip_add_mod_vs(uint64_t *result, const uint64_t *operand1, uint64_t operand2, uint64_t n,
                                      uint64_t modulus, void *stream);
/* EltwiseSubMod vector-vector: eltwise-sub-mod.hpp:13 (eltwise-sub-mod.cpp:15-30). */
axhel_hip_status axhel_hip_sub_mod_vv(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                                      uint64_t n, uint64_t modulus, void *stream);
/* EltwiseSubMod vector-scalar: eltwise-sub-mod.hpp:17 (eltwise-sub-mod.cpp:33-47). */
axhel_hip_status axhel_hip_sub_mod_vs(uint64_t *result, const uint64_t *operand1, uint64_t operand2, uint64_t n,
                                      uint64_t modulus, void *stream);
/* EltwiseMulMod: eltwise-mul-mod.hpp:14 (eltwise-mul-mod.cpp:19-94).
   input_mod_factor outside {1,2,4} -> AXHEL_HIP_UNSUPPORTED_FACTOR, no write. */
axhel_hip_status axhel_hip_mul_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2, uint64_t n,
                                   uint64_t modulus, uint64_t input_mod_factor, void *stream);
/* EltwiseFMAMod: eltwise-fma-mod.hpp:15 (eltwise-fma-mod.cpp:19-112).
   operand3 may be NULL.  Factor outside {1,2,4,8} -> UNSUPPORTED_FACTOR. */
axhel_hip_status axhel_hip_fma_mod(uint64_t *result, const uint64_t *operand1, uint64_t operand2,
                                   const uint64_t *operand3, uint64_t n, uint64_t modulus, uint64_t input_mod_factor,
                                   void *stream);
/* EltwiseReduceMod: eltwise-reduce-mod.hpp:14 (eltwise-reduce-mod.cpp:32-140).
   input_mod_factor == modulus selects the Barrett path for any 64-bit input.
   Combinations the reference skips -> UNSUPPORTED_FACTOR, no write. */
axhel_hip_status axhel_hip_reduce_mod(uint64_t *result, const uint64_t *operand, uint64_t n, uint64_t modulus,
                                      uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream);

/* -------------------------------------------------------------------- NTT */

/* An NTT plan: one (degree, modulus, root) with its twiddle tables resident
   in device memory.  Immutable after creation; safe to share across threads. */
typedef struct axhel_hip_plan axhel_hip_plan;

/* NTT::NTT(degree, modulus, root_of_unity): include/axhel/ntt/ntt.hpp:32
   (ntt.cpp:136-152, tables ntt-tables.cpp:81-140).  degree must be a power of
   two in [2, 2^20] with modulus == 1 mod 2*degree, modulus < 2^62. */
axhel_hip_status axhel_hip_ntt_plan_create(uint64_t degree, uint64_t modulus, uint64_t root_of_unity,
                                           axhel_hip_plan **plan);
/* NTT::NTT(degree, modulus, root_powers, inv_root_powers, inv_degree_modulo):
   ntt.hpp:37 (ntt.cpp:124-133).  Tables are host arrays of `degree`
   {operand, quotient} pairs (2*degree words, the NTTMultiplyOperand layout of
   ntt.hpp:14-18), borrowed for the plan's life as the reference borrows them:
   they must stay valid until the plan is destroyed.  The device holds a copy;
   axhel_hip_ntt_forward / _inverse compare the table they use with it and,
   if the caller has changed it, wait for the device to finish the work
   issued so far and reload it, so a call sees the tables as they are when it
   is made, as in the reference.  Batches, RNS contexts, products and the
   server use the device copy as last loaded. */
axhel_hip_status axhel_hip_ntt_plan_create_external(uint64_t degree, uint64_t modulus, const uint64_t *root_powers,
                                                    const uint64_t *inv_root_powers, uint64_t inv_degree_operand,
                                                    uint64_t inv_degree_quotient, axhel_hip_plan **plan);
axhel_hip_status axhel_hip_ntt_plan_destroy(axhel_hip_plan *plan);
/* NTT::Degree / NTT::Modulus: ntt.hpp:58-61 (ntt.cpp:210-216). */
uint64_t axhel_hip_ntt_plan_degree(const axhel_hip_plan *plan);
uint64_t axhel_hip_ntt_plan_modulus(const axhel_hip_plan *plan);
/* Copies the plan's host-side tables out (2*degree words each, plus
   {n^-1, quotient}); any pointer may be NULL. */
axhel_hip_status axhel_hip_ntt_plan_tables(const axhel_hip_plan *plan, uint64_t *root_powers,
                                           uint64_t *inv_root_powers, uint64_t *inv_degree);

/* NTT::ComputeForward: ntt.hpp:49 (ntt.cpp:173-186).  input_mod_factor is
   ignored as in the reference; output_mod_factor == 4 gives the lazy [0,4q)
   output, anything else the normalized [0,q) output.  Operand words are in
   [0,4q), the reference's domain (ntt-native.cpp:17). */
axhel_hip_status axhel_hip_ntt_forward(const axhel_hip_plan *plan, uint64_t *result, const uint64_t *operand,
                                       uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream);
/* NTT::ComputeInverse: ntt.hpp:53 (ntt.cpp:192-204).  output_mod_factor == 2
   gives the lazy [0,2q) output, anything else [0,q). */
axhel_hip_status axhel_hip_ntt_inverse(const axhel_hip_plan *plan, uint64_t *result, const uint64_t *operand,
                                       uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream);

/* ---------------------------------------------------------------- batching */

/* One limb-transform of a batch: what one ComputeForward/ComputeInverse call
   does, for a device-resident operand.  No reference counterpart: the
   reference handles one vector per call and leaves the limb loop to the
   caller (integration/seal/SEAL-4.4.5-AXHEL.patch:185-226). */
typedef struct axhel_hip_ntt_job {
    const axhel_hip_plan *plan;
    uint64_t *result;
    const uint64_t *operand;
    uint64_t output_mod_factor;
    int32_t inverse; /* 0 = ComputeForward, 1 = ComputeInverse */
    int32_t reserved;
} axhel_hip_ntt_job;

typedef struct axhel_hip_batch axhel_hip_batch;

/* Builds a reusable batch: jobs are grouped by (degree, direction), the job
   tables are uploaded to device memory, a kernel configuration is picked per
   degree (the autotuned table when one is loaded) and the launches are
   captured into one hipGraph.  Every job pointer must be device memory and no
   two jobs may write overlapping memory. */
axhel_hip_status axhel_hip_ntt_batch_create(const axhel_hip_ntt_job *jobs, size_t count, axhel_hip_batch **batch);
/* Runs the batch on `stream` (asynchronous): a batch of one (degree,
   direction) group as plain kernel launches of its passes, any other as a
   replay of the captured graph, whose parallel branches run the groups side
   by side. */
axhel_hip_status axhel_hip_ntt_batch_launch(axhel_hip_batch *batch, void *stream);
/* Replays the captured graph on `stream`, whatever the group count. */
axhel_hip_status axhel_hip_ntt_batch_launch_graph(axhel_hip_batch *batch, void *stream);
/* Same work as launch, issued as plain kernel launches without the graph. */
axhel_hip_status axhel_hip_ntt_batch_launch_eager(axhel_hip_batch *batch, void *stream);
axhel_hip_status axhel_hip_ntt_batch_destroy(axhel_hip_batch *batch);
/* Number of kernel launches one replay issues, and the index-th of them run
   alone on `stream`, in replay order (per-kernel timing with events). */
int axhel_hip_ntt_batch_kernel_count(const axhel_hip_batch *batch);
axhel_hip_status axhel_hip_ntt_batch_launch_kernel(axhel_hip_batch *batch, int index, void *stream);
/* Number of the batch's (degree, direction) groups whose ROW pass runs over
   stacks of jobs sharing one plan (one twiddle set per workgroup).  A group
   is stacked when its configuration has two passes and padding each plan's
   jobs to whole stacks leaves at most one slot in eight empty.  Diagnostic;
   no reference counterpart. */
int axhel_hip_ntt_batch_stacked_groups(const axhel_hip_batch *batch);
/* What the index-th launch of a replay computes (per-kernel rooflines).
   limbs counts the launch's limb-transforms (empty stack slots excluded);
   butterflies = (degree / 2) * log_bits * limbs; bytes = 16 * degree * limbs,
   every coefficient read once and written once by the pass.  schedule is
   the AXHEL_HIP_UBENCH_* butterfly schedule of the pass's stages
   (include/axhel_hip_ubench.h).  Diagnostic; no reference counterpart. */
typedef struct axhel_hip_ntt_kernel_info {
    uint64_t degree;
    uint32_t limbs;
    uint32_t pass;     /* 0 = the transform's first pass */
    uint32_t passes;   /* passes of the transform */
    uint32_t log_bits; /* index bits (butterfly stages) this pass covers */
    int32_t inverse;
    int32_t normalize; /* the pass folds the output normalization in */
    int32_t relaxed;   /* relaxed forward schedule */
    int32_t stacked;   /* ROW pass over stacks of same-plan jobs */
    int32_t schedule;
    int32_t reserved;
    uint64_t butterflies;
    uint64_t bytes;
} axhel_hip_ntt_kernel_info;
axhel_hip_status axhel_hip_ntt_batch_kernel_info(const axhel_hip_batch *batch, int index,
                                                 axhel_hip_ntt_kernel_info *info);
/* One-shot batch: create, launch eagerly on `stream`, synchronize, destroy. */
axhel_hip_status axhel_hip_ntt_batch_run(const axhel_hip_ntt_job *jobs, size_t count, void *stream);

/* ------------------------------------------------------ polynomial product */

/* One limb of an RNS polynomial product:
     result = InverseNTT(FMAMod(MulMod(NTT(a), NTT(b), 1), scalar, addend, 1))
   with normalized transforms: the call sequence of
   tests/public/test_polymul.cpp:13-25 (ComputeForward x2, EltwiseMulMod,
   ComputeInverse) with EltwiseFMAMod (eltwise-fma-mod.hpp:15) before the
   inverse.  Device pointers, N words each; addend may be NULL; scratch is N
   words the library may overwrite.  No reference counterpart as one call. */
typedef struct axhel_hip_polymul_job {
    const axhel_hip_plan *plan;
    uint64_t *result;
    const uint64_t *a;
    const uint64_t *b;
    const uint64_t *addend;
    uint64_t scalar;
    uint64_t *scratch;
} axhel_hip_polymul_job;

typedef struct axhel_hip_polymul axhel_hip_polymul;

/* fused != 0: forward COL passes of a and b, one fused kernel (forward ROW
   stages of both, MulMod, FMAMod, inverse ROW stages), inverse COL pass.
   fused == 0: the five calls as batched kernels, in the reference's order.
   Both give the same bits. */
axhel_hip_status axhel_hip_polymul_create(const axhel_hip_polymul_job *jobs, size_t count, int fused,
                                          axhel_hip_polymul **pm);
axhel_hip_status axhel_hip_polymul_launch(axhel_hip_polymul *pm, void *stream);
axhel_hip_status axhel_hip_polymul_destroy(axhel_hip_polymul *pm);
/* Number of the product's degree groups whose fused kernel runs over stacks
   of jobs sharing one plan (one LDS twiddle set per workgroup; same rule as
   axhel_hip_ntt_batch_stacked_groups).  Diagnostic; no reference
   counterpart. */
int axhel_hip_polymul_stacked_groups(const axhel_hip_polymul *pm);

/* ---------------------------------------------------------------- serving */

/* Request loop with continuous batching.  No reference counterpart: the
   reference is a synchronous single-call library; this is the caller loop
   (integration/seal/SEAL-4.4.5-AXHEL.patch:185-226 runs it per limb, one call
   at a time) moved onto the device.  A scheduler thread admits requests in
   arrival order while memory allows, packs every admitted request into the
   next device batch (one kernel launch per (op, degree) group), keeps up to
   `slots` batches in flight on separate streams, and completes requests as
   their batch finishes, so arrivals join while earlier work runs. */

typedef enum axhel_hip_op {
    AXHEL_HIP_OP_NTT_FORWARD = 0, /* ComputeForward per limb (plans[l]) */
    AXHEL_HIP_OP_NTT_INVERSE = 1, /* ComputeInverse per limb */
    AXHEL_HIP_OP_ADD = 2,         /* EltwiseAddMod vv per limb (moduli[l]) */
    AXHEL_HIP_OP_SUB = 3,         /* EltwiseSubMod vv */
    AXHEL_HIP_OP_MUL = 4,         /* EltwiseMulMod, input_mod_factor */
    AXHEL_HIP_OP_FMA = 5,         /* EltwiseFMAMod, scalars[l], operand2 = addend or NULL */
    AXHEL_HIP_OP_REDUCE = 6       /* EltwiseReduceMod arithmetic branches (2->1, 4->1, 4->2); input_mod_factor
                                     0 = each limb's Barrett branch (in_mod_factor == modulus) */
} axhel_hip_op;

typedef enum axhel_hip_backend {
    AXHEL_HIP_BACKEND_DEVICE = 0,
    /* Scheduler only: no GPU is touched and no result is computed; batches
       "finish" after null_latency_us.  For testing admission, timeouts,
       cancellation and fault injection on machines without a GPU. */
    AXHEL_HIP_BACKEND_NULL = 1
} axhel_hip_backend;

typedef struct axhel_hip_serve_config {
    uint64_t memory_budget_bytes; /* device bytes for staging host operands and the server's descriptor
                                     buffers; 0 = half the free HBM */
    uint32_t max_batch_limbs;     /* limbs packed into one device batch; 0 = 8192 */
    uint32_t slots;               /* device batches in flight; 0 = 4 */
    uint32_t backend;             /* axhel_hip_backend */
    uint32_t null_latency_us;     /* NULL backend: simulated batch duration */
    uint64_t fail_every_alloc;    /* fault injection: every k-th staging allocation fails; 0 = never */
    uint64_t fail_every_batch;    /* fault injection: every k-th device batch fails after its staging copies
                                     are queued (its requests end AXHEL_HIP_NO_MEMORY); 0 = never */
    uint64_t latency_samples;     /* latencies kept for p50/p99 (a uniform sample beyond); 0 = 2^20 */
} axhel_hip_serve_config;

/* One request: `limbs` vectors of `degree` words, limb l at offset l*degree
   in every buffer.  Buffers may be host or device memory; host buffers are
   staged through the server's device pool (and count against the budget).
   The arrays must stay valid until the request completes. */
typedef struct axhel_hip_request {
    int32_t op;    /* axhel_hip_op */
    uint32_t limbs;
    uint64_t degree;
    const axhel_hip_plan *const *plans; /* NTT ops: one plan per limb, of this degree */
    const uint64_t *moduli;             /* element-wise ops: one modulus per limb */
    const uint64_t *scalars;            /* FMA: one scalar per limb */
    uint64_t *result;
    const uint64_t *operand1;
    const uint64_t *operand2; /* ADD/SUB/MUL: second vector; FMA: addend (NULL allowed) */
    uint64_t input_mod_factor;
    uint64_t output_mod_factor;
    int64_t timeout_us; /* dropped with AXHEL_HIP_TIMEOUT if not started in time; <= 0: no limit */
} axhel_hip_request;

typedef struct axhel_hip_serve_stats {
    uint64_t submitted, completed, rejected, timed_out, cancelled, failed;
    uint64_t limbs_completed, batches;
    uint64_t bytes_in_use;      /* staging of admitted requests */
    uint64_t peak_bytes_in_use; /* peak of staging plus descriptor buffers */
    /* arrival -> completion of the requests completed since the last reset;
       p50/p99 exact up to latency_samples of them, then read from a uniform
       sample of that many (the server's memory stays bounded); mean and max
       exact */
    double p50_us, p99_us, mean_us, max_us;
    double first_arrival_to_last_completion_s;
    uint64_t descriptor_bytes; /* the slots' device descriptor buffers (charged to the budget) */
    /* device bytes of the staging pool, blocks in use and cached; cached
       blocks are freed before an allocation would take the pool plus the
       descriptor buffers past the budget */
    uint64_t pool_bytes, peak_pool_bytes;
} axhel_hip_serve_stats;

typedef struct axhel_hip_server axhel_hip_server;

axhel_hip_status axhel_hip_server_create(const axhel_hip_serve_config *config, axhel_hip_server **server);
/* Stops accepting requests, finishes every admitted one, fails the queued
   ones with AXHEL_HIP_CANCELLED, and frees the server. */
axhel_hip_status axhel_hip_server_destroy(axhel_hip_server *server);
/* Resident plan cache keyed by (degree, modulus, root), owned by the server. */
axhel_hip_status axhel_hip_server_plan(axhel_hip_server *server, uint64_t degree, uint64_t modulus,
                                       uint64_t root_of_unity, const axhel_hip_plan **plan);
/* Queues a request; *ticket identifies it.  Rejected at once (no ticket)
   when invalid or larger than the whole memory budget. */
axhel_hip_status axhel_hip_server_submit(axhel_hip_server *server, const axhel_hip_request *request,
                                         uint64_t *ticket);
/* Waits up to timeout_us (< 0: forever) for the request; on completion
   returns its final status (AXHEL_HIP_OK, or AXHEL_HIP_TIMEOUT for a request
   dropped at its deadline, AXHEL_HIP_CANCELLED, AXHEL_HIP_NO_MEMORY, ...) and
   forgets the ticket.  AXHEL_HIP_PENDING means the wait gave up first: the
   request is still queued or running and the ticket stays valid.  When
   several threads wait on one ticket, one of them collects the status and
   the others return AXHEL_HIP_INVALID_ARGUMENT. */
axhel_hip_status axhel_hip_server_wait(axhel_hip_server *server, uint64_t ticket, int64_t timeout_us);
/* Cancels a queued request (AXHEL_HIP_OK), or reports AXHEL_HIP_REJECTED if
   it already runs or finished. */
axhel_hip_status axhel_hip_server_cancel(axhel_hip_server *server, uint64_t ticket);
axhel_hip_status axhel_hip_server_stats(axhel_hip_server *server, axhel_hip_serve_stats *stats);
/* Clears latency samples and counters (after a warm-up). */
axhel_hip_status axhel_hip_server_reset_stats(axhel_hip_server *server);

/* ------------------------------------------------------------ RNS context */

/* N and L primes with one resident plan per prime, and calls over whole RNS
   polynomials: `count` polynomials stored contiguously as
   [count][limbs][degree] uint64 words in device memory (limb l of
   polynomial p at offset (p * limbs + l) * degree).  One call does what the
   reference's caller does with one ComputeForward/ComputeInverse or Eltwise*Mod
   call per limb (integration/seal/SEAL-4.4.5-AXHEL.patch:185-226: the limb
   loop); the arithmetic per limb is that call's.  Calls are asynchronous on
   `stream`; the context keeps one job-table buffer per stream, so calls may
   run on several streams at once.  Calls on a stream that is being captured
   into a graph are refused (AXHEL_HIP_INVALID_ARGUMENT): capture a batch
   instead.  No reference counterpart. */
typedef struct axhel_hip_rns axhel_hip_rns;

#define AXHEL_HIP_RNS_MAX_LIMBS 64
/* input_mod_factor of AXHEL_HIP_OP_REDUCE meaning "equal to each limb's
   modulus": Barrett reduction of arbitrary 64-bit words (the in == modulus
   branch of eltwise-reduce-mod.cpp:32-128). */
#define AXHEL_HIP_RNS_ANY_INPUT 0xFFFFFFFFFFFFFFFFull

/* roots: primitive 2N-th roots, one per modulus, or NULL for the first
   a = 2, 3, ... whose a^((q-1)/2N) has order 2N (parameters.hpp:110-122).
   limbs in [1, AXHEL_HIP_RNS_MAX_LIMBS]. */
axhel_hip_status axhel_hip_rns_create(uint64_t degree, const uint64_t *moduli, const uint64_t *roots, uint32_t limbs,
                                      axhel_hip_rns **ctx);
axhel_hip_status axhel_hip_rns_destroy(axhel_hip_rns *ctx);
/* The plan of limb l (owned by the context). */
const axhel_hip_plan *axhel_hip_rns_plan(const axhel_hip_rns *ctx, uint32_t limb);
/* Forward (inverse = 0) or inverse (1) NTT of every limb of `count`
   polynomials; output_mod_factor as ComputeForward/ComputeInverse (lazy for
   4 / 2, normalized otherwise).  result may equal operand. */
axhel_hip_status axhel_hip_rns_ntt(axhel_hip_rns *ctx, uint64_t *result, const uint64_t *operand, uint64_t count,
                                   int inverse, uint64_t output_mod_factor, void *stream);
/* axhel_hip_rns_ntt that allocates nothing.  No counterpart in the
   reference: it is for callers that cannot allocate device memory on the
   call path (code under graph capture elsewhere in the program, tight
   serving loops, harnesses that forbid hipMalloc).  The job and stack
   tables go to a fixed device array of AXHEL_HIP_RNS_NOALLOC_MAX_ENTRIES
   descriptors (1 MiB per GPU, held until the process ends) instead of a
   buffer the context allocates; a call needing more (count x limbs
   descriptors, doubled when the ROW pass runs stacked) returns
   AXHEL_HIP_INVALID_ARGUMENT and writes nothing.  The array's address is
   taken at axhel_hip_rns_create (or on the first call on a device).  Every
   call on one GPU uses the same array: calls that overlap on a GPU (other
   streams, other threads, other contexts) race and must be ordered by the
   caller.  Same results as axhel_hip_rns_ntt. */
#define AXHEL_HIP_RNS_NOALLOC_MAX_ENTRIES 16384
axhel_hip_status axhel_hip_rns_ntt_noalloc(axhel_hip_rns *ctx, uint64_t *result, const uint64_t *operand,
                                           uint64_t count, int inverse, uint64_t output_mod_factor, void *stream);
/* Element-wise op over every limb with that limb's modulus: op is
   AXHEL_HIP_OP_ADD / SUB (operand2, or NULL and scalars[limbs] on the host
   for the vector-scalar forms, eltwise-add-mod.hpp:17 and
   eltwise-sub-mod.hpp:17), MUL (operand2 required,
   input_mod_factor in {1,2,4}), FMA (scalars[limbs] on the host, operand2 =
   addend or NULL, input_mod_factor in {1,2,4,8}) or REDUCE (input/output
   factors as EltwiseReduceMod, AXHEL_HIP_RNS_ANY_INPUT for the Barrett
   branch).  Unsupported factors return AXHEL_HIP_UNSUPPORTED_FACTOR and leave
   result untouched. */
axhel_hip_status axhel_hip_rns_eltwise(axhel_hip_rns *ctx, int op, uint64_t *result, const uint64_t *operand1,
                                       const uint64_t *operand2, const uint64_t *scalars, uint64_t count,
                                       uint64_t input_mod_factor, uint64_t output_mod_factor, void *stream);

/* ---------------------------------------------------------------- tuning */

/* Kernel configuration for one degree.  A transform runs as one ROW pass
   (all index bits in one workgroup's LDS) or as a COL pass over the top
   log_col bits followed by a ROW pass over the rest (inverse: reversed).
   log_elems = log2 words per thread in the ROW pass, i.e. the butterfly
   stages done in registers between LDS exchanges (the radix); row_groups /
   col_groups = groups (rows / columns) per workgroup; log_elems_col = the
   COL pass's radix. */
typedef struct axhel_hip_ntt_config {
    int32_t log_elems;
    int32_t row_groups;
    int32_t col_groups;    /* 0 for a single pass */
    int32_t log_col;       /* 0 for a single pass */
    int32_t log_elems_col; /* 0 for a single pass */
} axhel_hip_ntt_config;

/* Number of compiled configurations for a degree, and the i-th of them. */
int axhel_hip_ntt_config_count(uint64_t degree);
axhel_hip_status axhel_hip_ntt_config_get(uint64_t degree, int index, axhel_hip_ntt_config *config);
/* Selects the configuration used for a degree from now on (process-wide),
   by batches and single calls alike; index -1 restores the tuning table's
   choices (ntt_tuning_gfx950.txt, or $AXHEL_HIP_TUNING; configuration 0 for
   a degree the table does not list). */
axhel_hip_status axhel_hip_ntt_config_select(uint64_t degree, int index);
int axhel_hip_ntt_config_selected(uint64_t degree);
/* The configuration single-limb calls (axhel_hip_ntt_forward / _inverse, the
   ComputeForward / ComputeInverse drop-ins) use: the tuning table's
   single-call entry for the degree (a shape chosen for latency), or the
   selected configuration once one has been selected. */
int axhel_hip_ntt_config_single(uint64_t degree);
/* Index of the configuration launches of 2 limbs up to 2^21 words (batch
   groups, RNS-context calls, serving groups) use: the tuning table's
   small-launch entry for the degree, else the single-call configuration.  No
   reference counterpart. */
int axhel_hip_ntt_config_small(uint64_t degree);

#ifdef __cplusplus
}
#endif

#endif /* AXHEL_HIP_H */
