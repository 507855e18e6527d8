"""Autotuner for the NTT kernel configurations.

For each degree it times every compiled configuration (radix = words per
thread, groups per workgroup, pass split) on a synthetic batch sized to fill
the GPU, forward and inverse, and keeps the fastest for batches; it also
times single-limb calls (the ComputeForward / ComputeInverse drop-ins, one
synchronous call at a time), whose latency favours other shapes, and keeps
the fastest of those for single calls, and times synchronous launches of a
few limbs for launches of 2 limbs up to 2^21 words.  The result is a small
text table, one `log2(N) batch-index single-index [small-index]` line per
degree (the last only where it differs from the single-call choice), that
the library reads at start-up (`axhel_amd/ntt_tuning_gfx950.txt`, or the
file named by AXHEL_HIP_TUNING).

Usage: python -m axhel_amd.autotune [--logn 12 13 ...] [--out PATH] [--words W]
"""
import argparse
import json
import os
import sys
import time

import axhel_amd as ax
from axhel_amd import params

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ntt_tuning_gfx950.txt")


def time_batch(batch, torch, reps, warm=10):
    """Seconds per replay, after `warm` untimed replays (the first dispatches
    after a host-side gap run at a lower clock)."""
    for _ in range(warm):
        batch.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune_degree(logn, torch, words, reps, limbs_per_prime=64):
    n = 1 << logn
    limbs = max(1, min(words // n, 16384))  # small degrees: a bounded job list, not words / n jobs
    nprimes = max(1, min(24, limbs // limbs_per_prime)) if limbs >= limbs_per_prime else 1
    primes = params.rns_primes(n, nprimes)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in primes]
    g = torch.Generator(device="cuda").manual_seed(logn)
    src = torch.empty((limbs, n), dtype=torch.int64, device="cuda")
    for i in range(limbs):
        src[i] = torch.randint(0, primes[i % nprimes], (n,), generator=g, device="cuda", dtype=torch.int64)
    dst = torch.empty_like(src)
    results = []
    for ci in range(ax.config_count(n)):
        ax.config_select(n, ci)
        fb = ax.Batch([(plans[i % nprimes], dst[i], src[i], False, 1) for i in range(limbs)])
        ib = ax.Batch([(plans[i % nprimes], src[i], dst[i], True, 1) for i in range(limbs)])
        tf = time_batch(fb, torch, reps)
        ti = time_batch(ib, torch, reps)
        fb.close()
        ib.close()
        results.append({"index": ci, "config": ax.config_get(n, ci), "fwd_ms": tf * 1e3, "inv_ms": ti * 1e3,
                        "limbs": limbs, "fwd_limbs_per_s": limbs / tf, "inv_limbs_per_s": limbs / ti})
        print(json.dumps({"logn": logn, **results[-1]}), flush=True)
    ax.config_select(n, -1)
    best = min(results, key=lambda r: r["fwd_ms"] + r["inv_ms"])
    return best["index"], results


def tune_single(logn, torch, calls=200):
    """Single-limb call latency per configuration: forward and inverse plain
    calls on device buffers, each followed by a synchronisation."""
    n = 1 << logn
    q = params.find_ntt_prime(n, 50)
    plan = ax.NTT(n, q, params.find_primitive_2n_root(n, q))
    src = torch.randint(0, q, (n,), dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    results = []
    for ci in range(ax.config_count(n)):
        ax.config_select(n, ci)
        us = 0.0
        for call in (plan.ComputeForward, plan.ComputeInverse):
            for _ in range(calls // 4):
                call(dst, src, 1, 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(calls):
                call(dst, src, 1, 1)
                torch.cuda.synchronize()
            us += (time.perf_counter() - t0) * 1e6 / calls
        results.append({"index": ci, "config": ax.config_get(n, ci), "fwd_plus_inv_us": us})
        print(json.dumps({"logn": logn, "single": True, **results[-1]}), flush=True)
    ax.config_select(n, -1)
    return min(results, key=lambda r: r["fwd_plus_inv_us"])["index"], results


def tune_small(logn, torch, calls=100):
    """Latency of a launch of a few limbs per configuration: synchronous
    batches (forward, then inverse) of half the most limbs that take the
    small-launch configuration (2^21 words), each limb with its own prime."""
    n = 1 << logn
    limbs = max(2, min(64, ((1 << 21) >> logn) // 2))
    primes = params.rns_primes(n, limbs)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in primes]
    src = torch.randint(0, min(primes), (limbs, n), dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    results = []
    for ci in range(ax.config_count(n)):
        ax.config_select(n, ci)
        us = 0.0
        for inverse in (False, True):
            b = ax.Batch([(plans[i], dst[i], src[i], inverse, 1) for i in range(limbs)])
            for _ in range(calls // 4):
                b.launch()
                torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(calls):
                b.launch()
                torch.cuda.synchronize()
            us += (time.perf_counter() - t0) * 1e6 / calls
            b.close()
        results.append({"index": ci, "config": ax.config_get(n, ci), "limbs": limbs, "fwd_plus_inv_us": us})
        print(json.dumps({"logn": logn, "small": True, **results[-1]}), flush=True)
    ax.config_select(n, -1)
    return min(results, key=lambda r: r["fwd_plus_inv_us"])["index"], results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--logn", type=int, nargs="*", default=list(range(10, 21)))
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--words", type=int, default=64 * 24 << 16,
                    help="coefficients per timed batch (default: the 64 x 24 x 2^16 RNS batch)")
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    import torch
    assert torch.cuda.is_available(), "autotune needs a GPU"
    # bring the clocks up before the first configuration is timed
    x = torch.randint(0, 1 << 40, (1 << 26,), dtype=torch.int64, device="cuda")
    t_end = time.perf_counter() + 1.0
    while time.perf_counter() < t_end:
        ax.EltwiseReduceMod(x, x, x.numel(), (1 << 50) - 27, 4, 2)
        torch.cuda.synchronize()
    del x
    header = ("# log2(N) batch-config-index single-call-config-index [small-launch-config-index], "
              "chosen by axhel_amd.autotune on gfx950\n")
    table, notes = {}, []
    if os.path.exists(args.out):
        for line in open(args.out):
            parts = line.split()
            if line.startswith("#"):
                if line != header:
                    notes.append(line)  # hand-written notes and overrides are kept
            elif len(parts) in (2, 3, 4):
                table[int(parts[0])] = [int(x) for x in parts[1:]]
    for logn in args.logn:
        best, _ = tune_degree(logn, torch, args.words, args.reps)
        single, _ = tune_single(logn, torch)
        small, _ = tune_small(logn, torch)
        table[logn] = [best, single] + ([small] if small != single else [])
    with open(args.out, "w") as f:
        f.write(header)
        f.writelines(notes)
        for logn in sorted(table):
            f.write(f"{logn} {' '.join(map(str, table[logn]))}\n")
    print("wrote", args.out, table, file=sys.stderr)


if __name__ == "__main__":
    main()
