"""Headline benchmark: NTT limbs/s at N = 2^16, L = 24 (batched vs plain).

Workload: 64 polynomials x 24 RNS limbs x N = 65536 uint64 coefficients
(1536 limb-transforms, 805 MB per operand array), 24 distinct primes of
50..60 bits (== 1 mod 2N), coefficients uniform in [0, q), resident in HBM.
One step = the forward negacyclic NTT (ComputeForward with default factors,
normalized output) of every limb, issued as one Batch.launch (a batch of one
degree and direction launches its two pass kernels directly; its captured
hipGraph is timed as a variant).  Plain baseline = ComputeForward per limb
through the same library, synchronous per call, on device-resident buffers.

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
    # one core, the reference benchmark's method (benchmark/src/axhel_benchmark.cpp:751-795):
    # 10 warm-up calls, then individually timed calls (at least 100),
    # Tukey-IQR filtered median and mean; tables built outside, in place
    for i in range(10):
        O.compute_forward_inplace(bufs[0], tabs[i % len(tabs)], 1)
    one_s = min(3.0, seconds / 4)
    samples, t1 = [], time.perf_counter()
    while len(samples) < 100 or time.perf_counter() - t1 < one_s:
        t = tabs[len(samples) % len(tabs)]
        a0 = time.perf_counter()
        O.compute_forward_inplace(bufs[0], t, 1)
        samples.append(time.perf_counter() - a0)
    x = np.array(samples)
    q1, q3 = np.percentile(x, [25, 75])
    kept = x[(x >= q1 - 1.5 * (q3 - q1)) & (x <= q3 + 1.5 * (q3 - q1))]
    med = float(np.median(kept))
    one_core = {"value": 1.0 / med, "unit": "limbs/s", "latency_us": med * 1e6,
                "latency_mean_us": float(kept.mean()) * 1e6,
                "sample": f"{len(x)} forward NTTs at N={n} on one thread after 10 warm-up calls; "
                          f"Tukey-IQR median of {len(kept)} kept samples"}
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


def polymul_leg(ax, torch, params, polys=128, limbs=32, logn=15, reps=20, warm=10, peaks=None):
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
    if peaks:
        # per product: NTT(a), NTT(b) (relaxed forward: COL stages + ROW stages
        # with the final stage and normalization) and one normalized inverse
        cfg = ax.config_get(n, ax.config_selected(n))
        col = cfg["log_col"]
        row = logn - col
        half = n // 2
        stages = {1: 2 * col, 2: 2 * row, 3: row, 5: col} if col else {2: 2 * logn, 5: logn}
        bfly = sum(stages.values()) * half * polys * limbs
        t_peak = sum(c * half * polys * limbs / (peaks[s]["peak"] * 1e9) for s, c in stages.items())
        ms = res["fused_ms"]
        ach = bfly / (ms * 1e-3) / 1e9
        passes_bytes = 80 * n * polys * limbs  # COL a, COL b, fused ROW (a, b, addend in; one out), inverse COL
        min_bytes = 32 * n * polys * limbs     # a, b, addend read once, result written once
        res["roofline"] = {
            "valu": {"achieved": ach, "peak": bfly / t_peak / 1e9, "unit": "Gbfly/s",
                     "frac": t_peak / (ms * 1e-3),
                     "peak_source": "schedule peaks weighted by the product's stages: "
                                    + ", ".join(f"{ax.SCHEDULE_NAMES[s]} x {c}" for s, c in stages.items())},
            "hbm": {"achieved": passes_bytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": passes_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "bytes_per_product": 80 * n, "minimum_bytes_per_product": 32 * n,
                    "frac_at_minimum_bytes": min_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    return res


def uneven_leg(ax, torch, params, jobs=512, reps=20, warm=10, seed=0, peaks=None):
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
    # residues uniform in [0, q) for each limb's own modulus
    src = torch.empty((total_words,), device="cuda", dtype=torch.int64)
    dst = torch.empty_like(src)
    g = torch.Generator(device="cuda").manual_seed(seed)
    blist = []
    off = 0
    limbs_total = 0
    for logn, limbs, inverse, k in spec:
        n = 1 << logn
        for l in range(limbs):
            plan = pools[logn][(k + l) % 40]
            src[off:off + n] = torch.randint(0, plan.modulus, (n,), generator=g, device="cuda", dtype=torch.int64)
            blist.append((plan, dst[off:off + n], src[off:off + n], inverse, 1))
            off += n
        limbs_total += limbs
    batch = ax.Batch(blist)
    for _ in range(warm):  # both modes warm before either is timed
        batch.launch_graph()
        batch.launch_eager()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch_graph()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch_eager()
    torch.cuda.synchronize()
    ems = (time.perf_counter() - t0) * 1e3 / reps
    infos = [batch.kernel_info(k) for k in range(batch.kernel_count())]
    batch.close()
    out = {"jobs": jobs, "limb_transforms": limbs_total, "words": total_words, "graph_ms": ms,
           "eager_ms": ems, "limbs_per_s": limbs_total / (ms * 1e-3),
           "coeff_per_s": total_words / (ms * 1e-3), "inputs": "uniform in [0, q) per limb"}
    if peaks:
        out["roofline"] = mix_roofline(infos, ms, peaks)
    return out


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


def sync_time(torch, fn, steps):
    """ms per call of fn over `steps` calls, bracketed by device synchronizations."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def steady_warmup(torch, fn, min_steps, block=5, tol=0.02, max_seconds=3.0):
    """Runs fn at least min_steps times, then in blocks of `block` calls until
    two consecutive blocks take the same time within `tol` (the clock has
    settled), at most max_seconds.  A cold GPU runs its first dispatches 2-4x
    slower while clocks ramp, so a fixed short warm-up leaves the timed steps
    on the ramp."""
    for _ in range(min_steps):
        fn()
    done = min_steps
    t_end = time.perf_counter() + max_seconds
    prev, last = None, None
    converged = False
    while time.perf_counter() < t_end:
        last = sync_time(torch, fn, block)
        done += block
        if prev is not None and abs(last - prev) <= tol * prev:
            converged = True
            break
        prev = last
    return {"requested": min_steps, "done": done, "converged": converged, "rule":
            f"at least the requested steps, then blocks of {block} until two consecutive blocks agree "
            f"within {tol:.0%} (at most {max_seconds:.0f} s)", "last_block_ms_per_step": last}


def instrumented_steps(torch, batch, steps, stream):
    """The timed steps once more, issued pass by pass with a HIP event
    recorded on the kernels' stream before each pass kernel and after the
    last: each launch's duration as it runs inside the step sequence (its
    neighbours the other passes, as in the timed region), averaged over the
    steps, and the wall time per step of this copy (to show the events do
    not change the step)."""
    nk = batch.kernel_count()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(nk + 1)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        for k in range(nk):
            evs[s][k].record(stream)
            batch.launch_kernel(k, stream)
        evs[s][nk].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    per = [float(np.mean([evs[s][k].elapsed_time(evs[s][k + 1]) for s in range(steps)])) for k in range(nk)]
    return per, wall


def _smi_sample():
    """Socket power (W) and the mean gfx clock over the XCDs (MHz) from
    amd-smi, or None when the tool is absent or fails."""
    import subprocess
    try:
        r = subprocess.run(["amd-smi", "metric", "-p", "-c", "-g", "0", "--json"], capture_output=True, text=True,
                           timeout=5)
        d = json.loads(r.stdout)
        g = d["gpu_data"][0] if isinstance(d, dict) else d[0]
        watts = float(g["power"]["socket_power"]["value"])
        clks = [float(v["clk"]["value"]) for k, v in g["clock"].items() if k.startswith("gfx_")]
        return watts, float(np.mean(clks))
    except Exception:
        return None


def power_leg(torch, work, seconds=1.5, skip=2):
    """Board power and gfx clock while `work` runs back to back for about
    `seconds` (a sampler thread calls amd-smi, ~5 samples/s; the first `skip`
    samples, taken while the clock ramps, are dropped): the median of each."""
    samples, stop = [], threading.Event()

    def sampler():
        while not stop.is_set():
            s = _smi_sample()
            if s is None:
                return
            samples.append(s)

    th = threading.Thread(target=sampler)
    th.start()
    t_end, k = time.perf_counter() + seconds, 0
    while time.perf_counter() < t_end:
        work()
        k += 1
        if k % 16 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    stop.set()
    th.join()
    kept = samples[skip:] if len(samples) > skip else samples
    if not kept:
        return None
    return {"watts": float(np.median([s[0] for s in kept])), "gfx_mhz": float(np.median([s[1] for s in kept])),
            "samples": len(kept)}


def schedule_peaks(ax, seconds=0.3):
    """Register-resident butterfly rate of every pass schedule: the
    microbenchmark back to back for ~`seconds`, median of the second half of
    the runs (the first runs after a host-side gap start at a lower clock and
    read ~10 % low; the first run is kept beside the peak)."""
    peaks = {}
    for s, name in enumerate(ax.SCHEDULE_NAMES):
        runs, secs, t_end = [], [], time.perf_counter() + seconds
        while time.perf_counter() < t_end or len(runs) < 8:
            r, sec = ax.butterfly_peak(s)
            runs.append(r / 1e9)
            secs.append(sec)
        peaks[s] = {"schedule": name, "peak": float(np.median(runs[len(runs) // 2:])),
                    "seconds": float(np.median(secs[len(secs) // 2:])), "first_run": runs[0], "runs": len(runs)}
    return peaks


def kernel_roofline(info, ms, peaks):
    """One pass launch against its bounds: butterflies/s over the
    register-resident peak of the same schedule, and algorithmic bytes/s over
    HBM peak.  The bound is the larger of the two fractions."""
    pk = peaks[info["schedule"]]
    bfly = info["butterflies"] / (ms * 1e-3) / 1e9
    gbs = info["bytes"] / (ms * 1e-3) / 1e9
    valu = {"achieved": bfly, "peak": pk["peak"], "unit": "Gbfly/s", "frac": bfly / pk["peak"],
            "peak_source": f"axhel_hip_ubench_schedule_peak({info['schedule']}) = {pk['schedule']}: "
                           f"median of the last {pk['runs'] - pk['runs'] // 2} of {pk['runs']} back-to-back runs",
            "peak_first_run": pk["first_run"]}
    hbm = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    bound = "valu" if valu["frac"] >= hbm["frac"] else "hbm"
    return {"ms": ms, "info": info, "valu": valu, "hbm": hbm, "bound": bound,
            "frac": max(valu["frac"], hbm["frac"])}


def valu_at_clock(kd, power):
    """The dominant pass's VALU fraction against its schedule's peak scaled
    to the gfx clock the board holds during the step: the microbenchmark runs
    at its own clock, while the step, with HBM streaming beside the
    arithmetic, may run lower when the board reaches its 1400 W limit."""
    if not power or "error" in power or not power.get("butterfly_peak"):
        return None
    f_step, f_peak = power["ntt_step"]["gfx_mhz"], power["butterfly_peak"]["gfx_mhz"]
    if f_step <= 0 or f_peak <= 0:
        return None
    peak = kd["valu"]["peak"] * f_step / f_peak
    return {"achieved": kd["valu"]["achieved"], "peak": peak, "unit": "Gbfly/s",
            "frac": kd["valu"]["achieved"] / peak, "step_gfx_mhz": f_step, "peak_gfx_mhz": f_peak,
            "step_watts": power["ntt_step"]["watts"], "peak_watts": power["butterfly_peak"]["watts"]}


def mix_roofline(infos, ms, peaks):
    """A whole replay of several launches (uneven degrees, both directions):
    butterflies/s against the time the same launches would take at their
    schedules' peaks, and algorithmic pass bytes/s against HBM peak."""
    bfly = sum(i["butterflies"] for i in infos)
    t_peak = sum(i["butterflies"] / (peaks[i["schedule"]]["peak"] * 1e9) for i in infos)
    byts = sum(i["bytes"] for i in infos)
    ach = bfly / (ms * 1e-3) / 1e9
    gbs = byts / (ms * 1e-3) / 1e9
    return {"valu": {"achieved": ach, "peak": bfly / t_peak / 1e9, "unit": "Gbfly/s", "frac": t_peak / (ms * 1e-3),
                     "peak_source": "schedule peaks weighted by each launch's butterflies"},
            "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                    "bytes": byts, "note": "16 N bytes per limb per pass"},
            "launches": len(infos)}


def batch_leg(ax, torch, batch, limbs, steps, warmup, stream, peaks):
    """A batch timed as Batch.launch calls on `stream` after a steady warm-up,
    with each launch's roofline from the instrumented copy of the steps."""
    wu = steady_warmup(torch, lambda: batch.launch(stream), warmup)
    ms = sync_time(torch, lambda: batch.launch(stream), steps)
    kd, iwall = instrumented_steps(torch, batch, steps, stream)
    kern = [kernel_roofline(batch.kernel_info(k), kd[k], peaks) for k in range(len(kd))]
    return {"ms_per_step": ms, "limbs_per_s": limbs / (ms * 1e-3), "warmup": wu, "kernels": kern,
            "instrumented_ms_per_step": iwall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
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
    ap.add_argument("--no-lazy", action="store_true")
    ap.add_argument("--no-power", action="store_true", help="skip the board power and clock samples")
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
    aux = torch.empty_like(src)
    torch.cuda.synchronize()

    jobs = [(plans[l], dst[p, l], src[p, l], False, 1) for p in range(P) for l in range(L)]
    batch = ax.Batch(jobs)
    limbs = P * L
    run_stream = torch.cuda.Stream()
    launch = lambda: batch.launch()

    # Warm-up on the headline step until its time settles, then the
    # variants (graph replays on the default and on a separate stream, eager
    # launches on the separate stream, one RNS-context call), then the
    # headline, Batch.launch() on the default stream (one group: its two pass
    # kernels launched directly), timed last so no variant rides the ramp.
    warm = steady_warmup(torch, launch, args.warmup)
    log(f"warm-up: {warm['done']} steps (converged {warm['converged']}, "
        f"last block {warm['last_block_ms_per_step']:.3f} ms/step)")
    graph_ms = sync_time(torch, lambda: batch.launch_graph(), args.steps)
    log(f"graph replay on the default stream: {graph_ms:.3f} ms/step")
    stream_ms = sync_time(torch, lambda: batch.launch_graph(run_stream), args.steps)
    log(f"graph replay on a separate stream: {stream_ms:.3f} ms/step")
    eager_ms = sync_time(torch, lambda: batch.launch_eager(run_stream), args.steps)
    log(f"same step as eager launches: {eager_ms:.3f} ms/step")
    # the same step as one RNS-context call (job tables built on the device
    # per call, no graph): what a SEAL-style limb loop gets without a Batch
    rns = ax.RNS(n, primes)
    rns_call = lambda: rns.ntt(dst, src, P, False, 1, stream=run_stream)
    for _ in range(3):
        rns_call()
    rns_ms = sync_time(torch, rns_call, args.steps)
    log(f"same step as one RNS-context call: {rns_ms:.3f} ms/step")
    rns.close()
    for _ in range(5):
        launch()
    # headline: exactly `steps` replays bracketed by device synchronizations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        launch()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    value = limbs * args.steps / el
    log(f"batched forward (Batch.launch, default stream): {ms:.3f} ms/step, {value:.4g} limbs/s")

    # per-kernel durations: the same steps again on the default stream with a
    # HIP event between the pass kernels; then the register-resident peak of
    # each pass schedule
    kdur, instr_ms = instrumented_steps(torch, batch, args.steps, torch.cuda.default_stream())
    peaks = schedule_peaks(ax)
    log(f"kernel avg ms: {[round(x, 4) for x in kdur]} (instrumented copy of the steps: {instr_ms:.3f} ms/step)")
    log("schedule peaks Gbfly/s:", {v["schedule"]: round(v["peak"], 1) for v in peaks.values()})
    kern = [kernel_roofline(batch.kernel_info(k), kdur[k], peaks) for k in range(len(kdur))]
    dom = int(np.argmax(kdur))
    kd = kern[dom]
    cfg = ax.config_get(n, ax.config_selected(n))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_n65536_64x24_hbm_bytes.txt")
    if os.path.exists(tpath) and args.logn == 16 and P == 64 and L == 24:
        for line in open(tpath):
            if line.startswith("traffic_bytes"):
                kv = dict(x.split("=") for x in line.split()[1:])
                traffic = float(kv.get(f"pass{dom}", "nan"))
    log(f"dominant pass {dom} ({kd['info']['schedule_name']}): {kd['valu']['achieved']:.1f} of "
        f"{kd['valu']['peak']:.1f} Gbfly/s = {kd['valu']['frac']:.3f}; {kd['hbm']['achieved']:.0f} GB/s = "
        f"{kd['hbm']['frac']:.3f} of HBM; bound {kd['bound']}")

    # board power and gfx clock: the headline step, the dominant pass's
    # butterfly microbenchmark (its peak's own clock) and a plain HBM stream
    power = None
    if not args.no_power:
        flat_s, flat_a = src.view(-1), aux.view(-1)
        words = flat_s.numel()
        power = {
            "ntt_step": power_leg(torch, launch),
            "butterfly_peak": power_leg(torch, lambda: ax.butterfly_peak(kd["info"]["schedule"])),
            "hbm_stream": power_leg(torch, lambda: ax.EltwiseReduceMod(flat_a, flat_s, words, 0x0FFFFFFFFFFC0001,
                                                                       4, 2)),
            "board_power_limit_w": 1400,
            "source": "amd-smi metric -p -c (socket power, mean gfx clock over the XCDs), median of the samples "
                      "taken while each workload runs back to back for ~1.5 s; 1400 W is MI355X's published "
                      "total board power",
        }
        if power["ntt_step"] is None:
            power = {"error": "amd-smi not available"}
        else:
            log("power:", json.dumps(power))

    # lazy transforms (SEAL's hot calls: ComputeForward(.., 4, 4) and
    # ComputeInverse(.., 2, 2)) and the normalized inverse, as batches
    legs = {}
    leg_specs = [("inverse", aux, dst, True, 1)]
    if not args.no_lazy:
        leg_specs += [("lazy_forward", aux, src, False, 4), ("lazy_inverse", aux, dst, True, 2)]
    for name, out_t, in_t, inverse, factor in leg_specs:
        b = ax.Batch([(plans[l], out_t[p, l], in_t[p, l], inverse, factor) for p in range(P) for l in range(L)])
        legs[name] = batch_leg(ax, torch, b, limbs, args.steps, args.warmup, run_stream, peaks)
        b.close()
        lk = legs[name]["kernels"]
        ld = int(np.argmax([k["ms"] for k in lk]))
        legs[name]["dominant"] = ld
        log(f"{name}: {legs[name]['ms_per_step']:.3f} ms/step, {legs[name]['limbs_per_s']:.4g} limbs/s; "
            f"dominant pass {ld} ({lk[ld]['info']['schedule_name']}) valu {lk[ld]['valu']['frac']:.3f} "
            f"hbm {lk[ld]['hbm']['frac']:.3f}")

    # VALU issue: wave64 VALU instructions per launch (SQ_INSTS_VALU of the
    # committed counter profile of this workload) over the launch duration
    valu_issue = None
    vpath = os.path.join(ROOT, "profiles", "ntt_n65536_64x24_valu_counters.txt")
    if os.path.exists(vpath) and args.logn == 16 and P == 64 and L == 24:
        insts = None
        for line in open(vpath):
            if line.startswith("valu_insts"):
                kv = dict(x.split("=") for x in line.split()[1:])
                insts = float(kv.get(f"pass{dom}", "nan"))
        if insts is not None:
            ach = insts / (kdur[dom] * 1e-3) / 1e9
            valu_issue = {"achieved": ach, "unit": "G wave64 VALU instr/s", "nominal": VALU_ISSUE_NOMINAL,
                          "frac_of_nominal": ach / VALU_ISSUE_NOMINAL, "insts_per_launch": insts,
                          "valu_per_butterfly": insts * 64.0 / kd["info"]["butterflies"]}

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

    del batch, aux
    pmul = None
    if not args.no_polymul:
        pmul = polymul_leg(ax, torch, params, peaks=peaks)
        log("polymul:", json.dumps(pmul))
    uneven = None
    if not args.no_uneven:
        uneven = uneven_leg(ax, torch, params, peaks=peaks)
        log("uneven batch:", json.dumps(uneven))

    serving = None
    if not args.no_serve:
        serving = serve_leg(ax, args.serve_seconds)
        log("serving:", json.dumps(serving)[:400])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, primes, psis, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    bound_obj = kd["valu"] if kd["bound"] == "valu" else kd["hbm"]
    out = {
        "metric": "NTT limbs/sec at N=2^16,L=24 (batched vs plain); p99 latency at fixed load",
        "value": value,
        "unit": "limbs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_done": warm,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: uniform residues in [0,q) per limb, generated on device (seed 0)",
        "config": {"workload": f"forward negacyclic NTT, {P} polys x {L} limbs x N={n}, "
                               "24 primes 50-60 bits, normalized output, one Batch.launch per step on the default stream "
                               "(two pass kernels launched directly)",
                   "polys": P, "limbs": L, "degree": n, "kernel_config": cfg},
        "plain_limbs_per_s": plain,
        "plain_host_buffers_limbs_per_s": plain_host,
        "batched_vs_plain": value / plain,
        "inverse_limbs_per_s": legs["inverse"]["limbs_per_s"],
        "lazy_forward_limbs_per_s": legs.get("lazy_forward", {}).get("limbs_per_s"),
        "lazy_inverse_limbs_per_s": legs.get("lazy_inverse", {}).get("limbs_per_s"),
        "kernel_ms": kdur,
        "kernel_timing": "HIP events on the default stream between the pass kernels of an instrumented copy of "
                         "the timed steps (the same batch and step count, issued pass by pass right after the "
                         "timed region), ms per launch averaged over the steps",
        "instrumented_ms_per_step": instr_ms,
        "kernels": kern,
        "legs": legs,
        "schedule_peaks": peaks,
        "eager_ms_per_step": eager_ms,
        "rns_context_ms_per_step": rns_ms,
        "graph_ms_per_step": graph_ms,
        "separate_stream_graph_ms_per_step": stream_ms,
        "roofline": {"bound": kd["bound"], "achieved": bound_obj["achieved"], "peak": bound_obj["peak"],
                     "unit": bound_obj["unit"], "frac": kd["frac"], "traffic": traffic,
                     "kernel": f"ntt pass {dom} of {len(kdur)} ({kdur[dom]:.4f} ms, "
                               f"{kd['info']['schedule_name']}, stacked={kd['info']['stacked']})",
                     "rule": "larger of the VALU fraction (butterflies/s over the register-resident peak of the "
                             "pass's own butterfly schedule) and the HBM fraction (16 N bytes per limb per pass "
                             "over 8 TB/s)",
                     "valu": kd["valu"], "hbm": kd["hbm"], "valu_issue": valu_issue,
                     "valu_at_step_clock": valu_at_clock(kd, power)},
        "power": power,
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
