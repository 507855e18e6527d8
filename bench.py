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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, primes, psis, seconds):
    """The oracle's forward NTT (C restatement of the reference Native path)
    on host threads, one independent limb per thread: limbs/s over a bounded
    sample of the same workload."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    tabs = [O.Tables(n, q, psi) for q, psi in zip(primes, psis)]
    rng = np.random.default_rng(7)
    bufs = [rng.integers(0, primes[i % len(primes)], size=n, dtype=np.uint64) for i in range(threads)]
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
            "sample": f"{total} forward NTTs at N={n} over {el:.1f} s, one limb per thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--polys", type=int, default=64)
    ap.add_argument("--limbs", type=int, default=24)
    ap.add_argument("--logn", type=int, default=16)
    ap.add_argument("--plain-limbs", type=int, default=384, help="limbs timed on the plain per-call path")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(limbs are independent: see the sharding note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import axhel_amd as ax
    from axhel_amd import params

    assert torch.cuda.is_available(), "bench needs a GPU"
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

    for _ in range(args.warmup):
        batch.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    value = limbs * args.steps / el
    log(f"batched forward: {ms:.3f} ms/step, {value:.4g} limbs/s")

    # inverse (reported beside)
    ibatch = ax.Batch([(plans[l], src[p, l], dst[p, l], True, 1) for p in range(P) for l in range(L)])
    ibatch.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ibatch.launch()
    torch.cuda.synchronize()
    inv_rate = limbs * args.steps / (time.perf_counter() - t0)
    # restore the forward inputs (inverse of forward is the identity)
    batch.launch()
    ibatch.launch()
    torch.cuda.synchronize()

    # per-kernel durations with HIP events on the stream the kernels run on
    stream = torch.cuda.Stream()
    kcount = batch.kernel_count()
    reps = max(3, min(10, args.steps))
    kdur = []
    with torch.cuda.stream(stream):
        for k in range(kcount):
            batch.launch_kernel(k, stream)  # warm
            evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
            for a, b in evs:
                a.record(stream)
                batch.launch_kernel(k, stream)
                b.record(stream)
            stream.synchronize()
            kdur.append(sum(a.elapsed_time(b) for a, b in evs) / reps)
    log("kernel avg ms:", [round(x, 4) for x in kdur])
    dom = int(np.argmax(kdur))
    # algorithmic bytes of one pass: every coefficient read once and written once
    alg_bytes = 16.0 * n * limbs
    achieved = alg_bytes / (kdur[dom] * 1e-3) / 1e9
    cfg = ax.config_get(n, ax.config_selected(n))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_n65536_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(f"pass{dom}_bytes")

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

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, primes, psis, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "NTT limbs/sec at N=2^16,L=24 (batched vs plain)",
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
                               "24 primes 50-60 bits, normalized output, one hipGraph replay per step",
                   "polys": P, "limbs": L, "degree": n, "kernel_config": cfg},
        "plain_limbs_per_s": plain,
        "batched_vs_plain": value / plain,
        "inverse_limbs_per_s": inv_rate,
        "kernel_ms": kdur,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"ntt pass {dom} of {kcount}"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
