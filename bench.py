"""Headline benchmark: NTT limbs/s at N = 2^16, L = 24 (batched vs plain).

Workload: 64 polynomials x 24 RNS limbs x N = 65536 uint64 coefficients
(1536 limb-transforms, 805 MB per operand array), 24 distinct primes of
50..60 bits (== 1 mod 2N), coefficients uniform in [0, q), resident in HBM.
One step = the forward negacyclic NTT (ComputeForward with default factors,
normalized output) of every limb, issued as one replay of the batch's
captured hipGraph.  Plain baseline = ComputeForward per limb through the same
library, synchronous per call, on device-resident buffers.

Prints one JSON line.  --gpus above 1 is refused: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# G wave64 VALU instructions/s if every instruction took 4 cycles per SIMD
# (256 CUs x 4 SIMDs, 2.4 GHz).  gfx950 issues some 32-bit ops faster: the
# butterfly microbenchmark itself exceeds this, so it is a nominal figure.
VALU_ISSUE_NOMINAL = 1024 * 2.4 / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, primes, psis, seconds):
    """The oracle's forward NTT (C restatement of the reference Native path)
    on host threads, one independent limb per thread: limbs/s over a bounded
    sample of the same workload.  A shorter single-thread sample first gives
    the one-core per-call latency."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    tabs = [O.Tables(n, q, psi) for q, psi in zip(primes, psis)]
    rng = np.random.default_rng(7)
    bufs = [rng.integers(0, primes[i % len(primes)], size=n, dtype=np.uint64) for i in range(threads)]
    one_s = min(3.0, seconds / 4)
    c1, t1 = 0, time.perf_counter()
    while time.perf_counter() - t1 < one_s:
        O.compute_forward_inplace(bufs[0], tabs[c1 % len(tabs)], 1)
        c1 += 1
    e1 = time.perf_counter() - t1
    one_core = {"value": c1 / e1, "unit": "limbs/s", "latency_us": e1 / c1 * 1e6,
                "sample": f"{c1} forward NTTs at N={n} on one thread over {e1:.1f} s"}
    counts = [0] * threads
    stop = time.perf_counter() + seconds

    def work(i):
        t = tabs[i % len(tabs)]
        a = bufs[i]
        c = 0
        while time.perf_counter() < stop:
            O.compute_forward_inplace(a, t, 1)
            c += 1
        counts[i] = c

    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    el = time.perf_counter() - t0
    total = sum(counts)
    return {"value": total / el, "unit": "limbs/s", "cores": threads, "kind": "port",
            "sample": f"{total} forward NTTs at N={n} over {el:.1f} s, one limb per thread",
            "host_cpu": _cpu_model(), "host_cpus_visible": os.cpu_count(), "one_core": one_core}


def eltwise_leg(ax, torch, src, dst, q, reps):
    """Each element-wise op over the whole resident array (one call, one
    modulus), timed with HIP events: algorithmic GB/s and fraction of HBM."""
    flat_a = src.view(-1)
    flat_r = dst.view(-1)
    n = flat_a.numel()
    half = n // 2
    a, b, r = flat_a[:half], flat_a[half:2 * half], flat_r[:half]
    st = torch.cuda.current_stream()
    ops = {
        "add_vv": (24, lambda: ax.EltwiseAddMod(r, a, b, half, q, stream=st)),
        "sub_vv": (24, lambda: ax.EltwiseSubMod(r, a, b, half, q, stream=st)),
        "mul_f1": (24, lambda: ax.EltwiseMulMod(r, a, b, half, q, 1, stream=st)),
        "fma_f1": (24, lambda: ax.EltwiseFMAMod(r, a, 12345, b, half, q, 1, stream=st)),
        "reduce_barrett": (16, lambda: ax.EltwiseReduceMod(r, a, half, q, q, 1, stream=st)),
        "reduce_4_1": (16, lambda: ax.EltwiseReduceMod(r, a, half, q, 4, 1, stream=st)),
    }
    out = {}
    st = torch.cuda.current_stream()
    for name, (bpe, fn) in ops.items():
        # reps launches back to back between one pair of events: the host's
        # launch latency hides behind the running kernels, so this is the
        # sustained duration of one launch
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = bpe * half / (ms * 1e-3) / 1e9
        out[name] = {"elements": half, "ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def polymul_leg(ax, torch, params, polys=128, limbs=32, logn=15, reps=20, warm=10):
    """Fused RNS polynomial product, 128 polys x 32 limbs x N=2^15: NTT(a),
    NTT(b), MulMod, FMAMod(scalar, addend), InverseNTT per limb; fused
    pipeline vs the same five calls as batched kernels."""
    n = 1 << logn
    qs = params.rns_primes(n, limbs)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]
    g = torch.Generator(device="cuda").manual_seed(3)
    shape = (polys, limbs, n)
    A = torch.empty(shape, dtype=torch.int64, device="cuda")
    B = torch.empty_like(A)
    C = torch.empty_like(A)
    for l, q in enumerate(qs):
        A[:, l] = torch.randint(0, q, (polys, n), generator=g, device="cuda", dtype=torch.int64)
        B[:, l] = torch.randint(0, q, (polys, n), generator=g, device="cuda", dtype=torch.int64)
        C[:, l] = torch.randint(0, q, (polys, n), generator=g, device="cuda", dtype=torch.int64)
    out = torch.empty_like(A)
    scr = torch.empty_like(A)
    res = {"workload": f"{polys} polys x {limbs} limbs x N={n}", "limb_products": polys * limbs}
    outs = {}
    for fused in (True, False):
        jobs = [(plans[l], out[p, l], A[p, l], B[p, l], C[p, l], 12345 + l, scr[p, l])
                for p in range(polys) for l in range(limbs)]
        pm = ax.Polymul(jobs, fused=fused)
        for _ in range(warm):  # the first dispatches of a cold GPU run slower while clocks ramp
            pm.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            pm.launch()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        key = "fused" if fused else "unfused"
        res[key + "_ms"] = ms
        res[key + "_limb_products_per_s"] = polys * limbs / (ms * 1e-3)
        outs[key] = out.clone()
        pm.close()
    res["fused_equals_unfused"] = bool(torch.equal(outs["fused"], outs["unfused"]))
    return res


def uneven_leg(ax, torch, params, jobs=512, reps=20, warm=10, seed=0):
    """512 jobs in one graph replay: N = 2^(12+U{0..5}), limbs U{1..40},
    moduli of 30..60 bits, forward and inverse mixed."""
    rng = np.random.default_rng(seed)
    pools = {}
    for logn in range(12, 18):
        n = 1 << logn
        qs, seen = [], {}
        for i in range(40):
            bits = 30 + (i * 31) // 40
            k = seen.get(bits, 0)
            seen[bits] = k + 1
            qs.append(params.ntt_primes(n, bits, k + 1)[k])
        pools[logn] = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]
    spec = []
    total_words = 0
    for j in range(jobs):
        logn = int(rng.integers(12, 18))
        limbs = int(rng.integers(1, 41))
        inverse = bool(rng.integers(0, 2))
        spec.append((logn, limbs, inverse, int(rng.integers(0, 40))))
        total_words += limbs << logn
    src = torch.randint(0, 1 << 29, (total_words,), device="cuda", dtype=torch.int64)
    dst = torch.empty_like(src)
    blist = []
    off = 0
    limbs_total = 0
    for logn, limbs, inverse, k in spec:
        n = 1 << logn
        for l in range(limbs):
            blist.append((pools[logn][(k + l) % 40], dst[off:off + n], src[off:off + n], inverse, 1))
            off += n
        limbs_total += limbs
    batch = ax.Batch(blist)
    for _ in range(warm):  # both modes warm before either is timed
        batch.launch()
        batch.launch_eager()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch_eager()
    torch.cuda.synchronize()
    ems = (time.perf_counter() - t0) * 1e3 / reps
    batch.close()
    return {"jobs": jobs, "limb_transforms": limbs_total, "words": total_words, "graph_ms": ms,
            "eager_ms": ems, "limbs_per_s": limbs_total / (ms * 1e-3),
            "coeff_per_s": total_words / (ms * 1e-3)}


def serve_leg(ax, seconds):
    """Serving loop under Poisson load (native driver axhel_serve_bench):
    capacity from a closed loop, then p50/p99 at 0.5x and 0.8x of it."""
    import subprocess
    exe = os.path.join(ax.LIB_DIR, "axhel_serve_bench")
    r = subprocess.run([exe, "--seconds", str(seconds), "--fractions", "0.5,0.8"], capture_output=True, text=True,
                       timeout=120 + 4 * seconds)
    if r.returncode != 0:
        log("serve bench failed:", r.stderr[-2000:])
        return {"error": r.returncode}
    return json.loads(r.stdout.strip().splitlines()[-1])


def p99_summary(serving):
    """The serving loop's p99 at each fixed offered load (Poisson arrivals at
    a stated fraction of the measured capacity), lifted beside the headline."""
    if not serving or "loads" not in serving:
        return None
    return [{"offered_req_per_s": x["offered_req_per_s"], "fraction_of_capacity": x["fraction"],
             "p50_us": x["p50_us"], "p99_us": x["p99_us"], "mix": serving.get("mix")} for x in serving["loads"]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--polys", type=int, default=64)
    ap.add_argument("--limbs", type=int, default=24)
    ap.add_argument("--logn", type=int, default=16)
    ap.add_argument("--plain-limbs", type=int, default=384, help="limbs timed on the plain per-call path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eltwise", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-polymul", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run one 2^27-word EltwiseReduceMod(4->2) (1 GiB in, 1 GiB out): "
                         "the known-size stream that scales rocprof FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(RNS limbs are independent: N GPUs would each run this batch over their share of the primes)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import axhel_amd as ax
    from axhel_amd import params

    assert torch.cuda.is_available(), "bench needs a GPU"
    if args.calibrate_copy:
        words = 1 << 27
        a = torch.zeros(words, dtype=torch.int64, device="cuda")
        b = torch.empty_like(a)
        ax.EltwiseReduceMod(b, a, words, 0x0FFFFFFFFFFC0001, 4, 2)
        torch.cuda.synchronize()
        log(f"copy calibration: {words} words read and written")
        return
    n, L, P = 1 << args.logn, args.limbs, args.polys
    primes = params.rns_primes(n, L)
    psis = [params.find_primitive_2n_root(n, q) for q in primes]
    plans = [ax.NTT(n, q, psi) for q, psi in zip(primes, psis)]
    log(f"workload: {P} polys x {L} limbs x N={n}")

    # inputs resident in HBM: uniform in [0, q) per limb, generated on device
    g = torch.Generator(device="cuda").manual_seed(0)
    src = torch.empty((P, L, n), dtype=torch.int64, device="cuda")
    for l, q in enumerate(primes):
        src[:, l, :] = torch.randint(0, q, (P, n), generator=g, device="cuda", dtype=torch.int64)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()

    jobs = [(plans[l], dst[p, l], src[p, l], False, 1) for p in range(P) for l in range(L)]
    batch = ax.Batch(jobs)
    limbs = P * L

    # Timed steps: graph replays on the default stream, as Batch.launch()
    # issues them.  Beside it: the same replays on a separate stream, and the
    # same kernels as plain launches (no graph) on that stream.
    run_stream = torch.cuda.Stream()
    for _ in range(args.warmup):
        batch.launch()
        batch.launch(run_stream)
        batch.launch_eager(run_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    value = limbs * args.steps / el
    log(f"batched forward (graph replay, default stream): {ms:.3f} ms/step, {value:.4g} limbs/s")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch(run_stream)
    torch.cuda.synchronize()
    stream_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"graph replay on a separate stream: {stream_ms:.3f} ms/step")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch_eager(run_stream)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"same step as eager launches: {eager_ms:.3f} ms/step")

    # the same step as one RNS-context call (job tables built on the device
    # per call, no graph): what a SEAL-style limb loop gets without a Batch
    rns = ax.RNS(n, primes)
    for _ in range(10):
        rns.ntt(dst, src, P, False, 1, stream=run_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rns.ntt(dst, src, P, False, 1, stream=run_stream)
    torch.cuda.synchronize()
    rns_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"same step as one RNS-context call: {rns_ms:.3f} ms/step")
    rns.close()

    # inverse (reported beside)
    ibatch = ax.Batch([(plans[l], src[p, l], dst[p, l], True, 1) for p in range(P) for l in range(L)])
    ibatch.launch(run_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ibatch.launch(run_stream)
    torch.cuda.synchronize()
    inv_rate = limbs * args.steps / (time.perf_counter() - t0)
    # restore the forward inputs (inverse of forward is the identity)
    batch.launch()
    ibatch.launch()
    torch.cuda.synchronize()

    # per-kernel durations with HIP events on the stream the kernels run on
    stream = torch.cuda.Stream()
    kcount = batch.kernel_count()
    reps = max(5, min(20, args.steps))
    kdur = []
    with torch.cuda.stream(stream):
        for k in range(kcount):
            # reps launches of kernel k back to back between one pair of
            # events on its stream (sustained duration of one launch)
            batch.launch_kernel(k, stream)  # warm
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(reps):
                batch.launch_kernel(k, stream)
            b.record(stream)
            stream.synchronize()
            kdur.append(a.elapsed_time(b) / reps)
    log("kernel avg ms:", [round(x, 4) for x in kdur])
    dom = int(np.argmax(kdur))
    # Algorithmic work of one pass launch: each of the two passes runs half of
    # the log2(N) stages, (N/2)*(log2(N)/2) butterflies per limb, and reads and
    # writes every coefficient once (16 N bytes per limb).
    bfly = (n // 2) * (args.logn / 2.0) * limbs
    alg_bytes = 16.0 * n * limbs
    achieved_bfly = bfly / (kdur[dom] * 1e-3) / 1e9
    achieved_gbs = alg_bytes / (kdur[dom] * 1e-3) / 1e9
    # register-resident butterfly peak: the microbenchmark back to back for
    # ~0.3 s, median of the second half.  The first runs after the host-side
    # gap start at a lower clock and read ~10 % low; the first run is
    # reported beside the peak.
    runs, secs, t_end = [], [], time.perf_counter() + 0.3
    while time.perf_counter() < t_end or len(runs) < 8:
        r, sec = ax.butterfly_peak()
        runs.append(r / 1e9)
        secs.append(sec)
    peak_bfly = float(np.median(runs[len(runs) // 2:]))
    peak_secs = float(np.median(secs[len(secs) // 2:]))
    cfg = ax.config_get(n, ax.config_selected(n))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_n65536_64x24_hbm_bytes.txt")
    if os.path.exists(tpath) and args.logn == 16 and P == 64 and L == 24:
        for line in open(tpath):
            if line.startswith("traffic_bytes"):
                kv = dict(x.split("=") for x in line.split()[1:])
                traffic = float(kv.get(f"pass{dom}", "nan"))
    # VALU issue: wave64 VALU instructions per launch (SQ_INSTS_VALU of the
    # committed counter profile of this workload) over the launch duration,
    # against the rate the butterfly microbenchmark sustains (its own
    # SQ_INSTS_VALU per launch from the same profile over its warmed launch
    # duration); the nominal 4-cycle figure is reported beside it.
    valu_issue = None
    vpath = os.path.join(ROOT, "profiles", "ntt_n65536_64x24_valu_counters.txt")
    if os.path.exists(vpath) and args.logn == 16 and P == 64 and L == 24:
        insts = ub_insts = None
        for line in open(vpath):
            f = line.split()
            if line.startswith("valu_insts"):
                kv = dict(x.split("=") for x in f[1:])
                insts = float(kv.get(f"pass{dom}", "nan"))
            elif len(f) == 4 and f[0] == "bfly_peak_kernel" and f[2] == "SQ_INSTS_VALU":
                ub_insts = float(f[3])
        if insts is not None and ub_insts is not None:
            ach = insts / (kdur[dom] * 1e-3) / 1e9
            # each microbenchmark call makes a 64-iteration warm-up launch and
            # the timed 2048-iteration launch of the same grid; the profile
            # averages the two, so the timed launch holds 2 / (1 + 64/2048)
            # of that average
            ub_timed = ub_insts * 2.0 / (1.0 + 64.0 / 2048.0)
            ub_rate = ub_timed / peak_secs / 1e9
            valu_issue = {"achieved": ach, "peak": ub_rate, "unit": "G wave64 VALU instr/s", "frac": ach / ub_rate,
                          "peak_source": "butterfly microbenchmark's SQ_INSTS_VALU per timed launch over its warmed launch time",
                          "ubench_valu_per_butterfly": ub_timed * 64.0 / (peak_bfly * 1e9 * peak_secs),
                          "nominal": VALU_ISSUE_NOMINAL, "frac_of_nominal": ach / VALU_ISSUE_NOMINAL,
                          "insts_per_launch": insts}
    log(f"dominant pass {dom}: {achieved_bfly:.1f} Gbfly/s of {peak_bfly:.1f} peak "
        f"(first of {len(runs)} runs {runs[0]:.1f}); {achieved_gbs:.0f} GB/s"
        + (f"; VALU issue {valu_issue['frac']:.2f} of peak" if valu_issue else ""))

    # plain per-call path: one synchronous ComputeForward per limb
    pl = min(args.plain_limbs, limbs)
    for i in range(8):
        plans[i % L].ComputeForward(dst[0, i % L], src[0, i % L])
        ax.synchronize()
    t0 = time.perf_counter()
    for i in range(pl):
        p, l = divmod(i, L)
        plans[l].ComputeForward(dst[p, l], src[p, l])
        ax.synchronize()
    plain = pl / (time.perf_counter() - t0)
    log(f"plain per-call: {plain:.4g} limbs/s ({pl} limbs)")
    # the same call on host buffers: the library stages them through device
    # memory (PCIe both ways), as a reference caller with host vectors gets
    hl = min(32, pl)
    hsrc = [src[0, l % L].cpu().numpy().view(np.uint64).copy() for l in range(hl)]
    hdst = np.empty(n, dtype=np.uint64)
    plans[0].ComputeForward(hdst, hsrc[0])
    t0 = time.perf_counter()
    for i in range(hl):
        plans[i % L].ComputeForward(hdst, hsrc[i])
    plain_host = hl / (time.perf_counter() - t0)
    log(f"plain per-call, host buffers (PCIe-inclusive): {plain_host:.4g} limbs/s ({hl} limbs)")

    elt = None
    if not args.no_eltwise:
        elt = eltwise_leg(ax, torch, src, dst, primes[0], reps=max(5, min(20, args.steps)))
        log("eltwise GB/s:", {k: round(v["GB/s"], 1) for k, v in elt.items()})

    del batch, ibatch
    pmul = None
    if not args.no_polymul:
        pmul = polymul_leg(ax, torch, params)
        log("polymul:", json.dumps(pmul))
    uneven = None
    if not args.no_uneven:
        uneven = uneven_leg(ax, torch, params)
        log("uneven batch:", json.dumps(uneven))

    serving = None
    if not args.no_serve:
        serving = serve_leg(ax, args.serve_seconds)
        log("serving:", json.dumps(serving)[:400])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, primes, psis, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "NTT limbs/sec at N=2^16,L=24 (batched vs plain); p99 latency at fixed load",
        "value": value,
        "unit": "limbs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: uniform residues in [0,q) per limb, generated on device (seed 0)",
        "config": {"workload": f"forward negacyclic NTT, {P} polys x {L} limbs x N={n}, "
                               "24 primes 50-60 bits, normalized output, one hipGraph replay per step on the default stream",
                   "polys": P, "limbs": L, "degree": n, "kernel_config": cfg},
        "plain_limbs_per_s": plain,
        "plain_host_buffers_limbs_per_s": plain_host,
        "batched_vs_plain": value / plain,
        "inverse_limbs_per_s": inv_rate,
        "kernel_ms": kdur,
        "eager_ms_per_step": eager_ms,
        "rns_context_ms_per_step": rns_ms,
        "separate_stream_graph_ms_per_step": stream_ms,
        "roofline": {"bound": "valu", "achieved": achieved_bfly, "peak": peak_bfly, "unit": "Gbfly/s",
                     "frac": achieved_bfly / peak_bfly, "traffic": traffic,
                     "kernel": f"ntt pass {dom} of {kcount} ({kdur[dom]:.4f} ms)",
                     "peak_source": "register-resident Harvey butterflies, all CUs (axhel_hip_ubench_butterfly_peak), "
                                    f"median of the last {len(runs) - len(runs) // 2} of {len(runs)} back-to-back runs",
                     "peak_first_run": runs[0],
                     "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved_gbs / HBM_PEAK_GBS},
                     "valu_issue": valu_issue},
        "eltwise": elt,
        "polymul": pmul,
        "uneven_batch": uneven,
        "serving": serving,
        "p99_latency": p99_summary(serving),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
