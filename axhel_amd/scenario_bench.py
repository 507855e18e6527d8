"""Scenario benchmark of the drop-in API in the reference benchmark's report
format (benchmark/src/axhel_benchmark.cpp, benchmark/README.md).

The same matrix -- kernels x polynomial degrees x modulus bit widths -- with
the same deterministic moduli, roots and operands as the reference's
benchmark, measured per public-API invocation on the GPU and written as
`summary.csv`, `raw_samples.csv` and `metadata.txt` with the reference's
columns, so GPU rows sit beside the CPU backends' rows:

    python -m axhel_amd.scenario_bench --output-dir results/gpu
    python -m axhel_amd.scenario_bench --compare results/gpu/summary.csv \\
        cpu/native/summary.csv --cpu-ghz 3.0 --output speedup_gpu.csv

The reference counts CPU cycles; here the `*_cycles` columns hold
nanoseconds per invocation (`metadata.txt` says so): device time between
two HIP events on the call's stream (`--timer device`, default), or host
time around the call and a stream synchronisation (`--timer host`, what a
synchronous C++ caller waits).  `--compare` converts a CPU summary's cycles
to nanoseconds with the given clock and reports CPU/GPU ratios.

Operands live in device memory and are generated on the host exactly as the
reference does (mt19937_64 seeded per scenario and kernel, rejection-sampled
bounded draws), then copied once.  NTTs run in place on a work buffer that
is restored from the input before every sample, outside the timed interval
(axhel_benchmark.cpp:1140-1167).
"""
import argparse
import csv
import datetime
import os
import platform
import random
import sys
import time

import numpy as np

from . import params

# axhel_benchmark.cpp:470-478
KERNELS = ["add_vv", "add_vs", "sub_vv", "sub_vs", "mul_f1", "mul_f2", "mul_f4", "fma_f1", "fma_f2", "fma_f4",
           "fma_f8", "reduce_2to1", "reduce_4to1", "reduce_4to2", "reduce_barrett", "ntt_forward",
           "ntt_forward_lazy", "ntt_inverse", "ntt_inverse_lazy"]
DEFAULT_DEGREES = [2048, 4096, 8192, 16384, 32768]                 # axhel_benchmark.cpp:49
DEFAULT_BITS = [20, 25, 30, 40, 50, 52, 54, 56, 58, 60, 61, 62]     # axhel_benchmark.cpp:50
DEFAULT_SEED = 0x415848454C42454E                                  # axhel_benchmark.cpp:46 ("AXHELBEN")
BACKEND = "HIP-gfx950"

# axhel_benchmark.cpp:977-986 (column order of summary.csv and raw_samples.csv)
SUMMARY_COLUMNS = ["backend", "kernel", "poly_degree", "qi_bits", "modulus", "mod_factor", "input_mod_factor",
                   "output_mod_factor", "samples_total", "samples_kept", "outliers_removed", "raw_min_cycles",
                   "raw_max_cycles", "q1_cycles", "q3_cycles", "iqr_cycles", "lower_fence_cycles",
                   "upper_fence_cycles", "min_cycles", "max_cycles", "mean_cycles", "median_cycles",
                   "stddev_cycles", "cv_percent", "mean_cycles_per_coeff", "median_cycles_per_coeff", "perf_scaled"]
RAW_COLUMNS = ["backend", "kernel", "poly_degree", "qi_bits", "modulus", "mod_factor", "input_mod_factor",
               "output_mod_factor", "repetition", "batch", "pmu_profile", "raw_cycles", "time_enabled",
               "time_running", "cycles", "perf_scaled", "cntvct_ticks", "instructions", "ipc", "l1d_cache_rd",
               "l1d_cache_refill", "l2d_cache_rd", "l2d_cache_refill", "stall_backend", "stall_backend_mem",
               "stall_frontend"]

M64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (the C++ standard's 64-bit Mersenne Twister), with the
    state twist and tempering vectorised over numpy arrays."""

    N, M = 312, 156
    MATRIX_A = np.uint64(0xB5026F5AA96619E9)
    UPPER = np.uint64(0xFFFFFFFF80000000)
    LOWER = np.uint64(0x000000007FFFFFFF)

    def __init__(self, seed):
        mt = [seed & M64]
        for i in range(1, self.N):
            prev = mt[-1]
            mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & M64)
        self.mt = np.array(mt, dtype=np.uint64)
        self.buf = np.empty(0, dtype=np.uint64)

    def _twist(self):
        mt, N, M = self.mt, self.N, self.M
        one = np.uint64(1)

        def step(cur, nxt, far):
            y = (cur & self.UPPER) | (nxt & self.LOWER)
            return far ^ (y >> one) ^ np.where((y & one) != 0, self.MATRIX_A, np.uint64(0))

        # entries 0..M-1 read only old words; M..N-2 read new words 0..N-M-2
        # and old neighbours; N-1 reads new words 0 and M-1
        mt[:N - M] = step(mt[:N - M], mt[1:N - M + 1], mt[M:N])
        mt[N - M:N - 1] = step(mt[N - M:N - 1], mt[N - M + 1:N], mt[:M - 1])
        mt[N - 1:N] = step(mt[N - 1:N], mt[0:1], mt[M - 1:M])
        z = mt.copy()
        z ^= (z >> np.uint64(29)) & np.uint64(0x5555555555555555)
        z ^= (z << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
        z ^= (z << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
        z ^= z >> np.uint64(43)
        return z

    def draws(self, count):
        """The next `count` raw 64-bit outputs."""
        out = [self.buf]
        have = self.buf.size
        while have < count:
            z = self._twist()
            out.append(z)
            have += z.size
        allv = np.concatenate(out)
        self.buf = allv[count:]
        return allv[:count]

    def next(self):
        return int(self.draws(1)[0])

    def bounded(self, n, bound):
        """n draws of bounded_random(rng, bound): rejection below
        (2^64 - bound) mod bound, then x mod bound (axhel_benchmark.cpp:709-716)."""
        threshold = np.uint64(((1 << 64) - bound) % bound)
        b = np.uint64(bound)
        res = []
        need = n
        while need > 0:
            x = self.draws(need)
            ok = x >= threshold
            kept = np.flatnonzero(ok)
            if kept.size >= need:
                cut = kept[need - 1] + 1
                # draws past the last accepted one go back to the stream
                self.buf = np.concatenate([x[cut:], self.buf])
                res.append(x[:cut][ok[:cut]] % b)
                need = 0
            else:
                res.append(x[ok] % b)
                need -= kept.size
        return np.concatenate(res) if res else np.empty(0, dtype=np.uint64)

    def factor_values(self, n, q, factor):
        """random_vector(n, q, factor): bounded draws below factor*q, or raw
        draws when factor*q reaches 2^64 (axhel_benchmark.cpp:718-730)."""
        if q * factor > M64:
            return self.draws(n).copy()
        return self.bounded(n, q * factor)


def scenario_seed(seed, degree, bits, kernel):
    """Per-scenario operand seed (axhel_benchmark.cpp:732-739)."""
    x = (seed ^ ((degree + 0x9E3779B97F4A7C15 + (seed << 6) + (seed >> 2)) & M64)) & M64
    x ^= (bits * 0xBF58476D1CE4E5B9) & M64
    for c in kernel.encode():
        x ^= (c + 0x9E3779B97F4A7C15 + ((x << 6) & M64) + (x >> 2)) & M64
        x &= M64
    return x


def quantile_sorted(v, p):
    """Linear interpolation over sorted samples (axhel_benchmark.cpp:741-749)."""
    if not v:
        raise ValueError("quantile on empty data")
    if len(v) == 1:
        return v[0]
    pos = p * (len(v) - 1)
    lo, hi = int(np.floor(pos)), int(np.ceil(pos))
    return v[lo] + (pos - lo) * (v[hi] - v[lo])


def statistics(samples):
    """Tukey IQR filter and summary statistics (axhel_benchmark.cpp:751-795)."""
    v = sorted(float(x) for x in samples)
    if not v:
        raise ValueError("no samples")
    st = {"samples_total": len(v), "raw_min": v[0], "raw_max": v[-1],
          "q1": quantile_sorted(v, 0.25), "q3": quantile_sorted(v, 0.75)}
    st["iqr"] = st["q3"] - st["q1"]
    st["lower_fence"] = st["q1"] - 1.5 * st["iqr"]
    st["upper_fence"] = st["q3"] + 1.5 * st["iqr"]
    kept = [x for x in v if st["lower_fence"] <= x <= st["upper_fence"]]
    if not kept:
        raise RuntimeError("IQR filter removed every sample")
    st["samples_kept"] = len(kept)
    st["outliers_removed"] = len(v) - len(kept)
    st["min"], st["max"] = kept[0], kept[-1]
    st["mean"] = sum(kept) / len(kept)
    st["median"] = quantile_sorted(kept, 0.5)
    ss = sum((x - st["mean"]) ** 2 for x in kept)
    st["stddev"] = (ss / (len(kept) - 1)) ** 0.5 if len(kept) > 1 else 0.0
    st["cv_percent"] = 100.0 * st["stddev"] / st["mean"] if st["mean"] != 0.0 else 0.0
    return st


def kernel_meta(name, q):
    """(mod_factor, input_mod_factor, output_mod_factor) as the reference's
    report rows carry them (axhel_benchmark.cpp:1056-1167)."""
    if name.startswith("mul_f") or name.startswith("fma_f"):
        return int(name[5:]), 0, 0
    red = {"reduce_2to1": (2, 1), "reduce_4to1": (4, 1), "reduce_4to2": (4, 2), "reduce_barrett": (q, 1)}
    if name in red:
        return (0,) + red[name]
    ntt = {"ntt_forward": (1, 1), "ntt_forward_lazy": (1, 4), "ntt_inverse": (1, 1), "ntt_inverse_lazy": (1, 2)}
    if name in ntt:
        return (0,) + ntt[name]
    return 0, 0, 0


class Timer:
    """Nanoseconds per sample of `count` back-to-back invocations."""

    def __init__(self, torch, mode):
        self.torch, self.mode = torch, mode
        self.stream = torch.cuda.current_stream()

    def __call__(self, invoke, count):
        torch = self.torch
        if self.mode == "host":
            torch.cuda.synchronize()
            t0 = time.perf_counter_ns()
            for _ in range(count):
                invoke()
            self.stream.synchronize()
            return float(time.perf_counter_ns() - t0)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(self.stream)
        for _ in range(count):
            invoke()
        b.record(self.stream)
        b.synchronize()
        return a.elapsed_time(b) * 1e6


def measure(timer, warmup, repetitions, batch, prepare, invoke):
    """Warm-up, then `repetitions` samples of `batch` invocations each,
    `prepare` outside every timed interval (axhel_benchmark.cpp:797-840).
    Returns (total ns, ns per invocation) per sample."""
    for _ in range(warmup):
        prepare()
        for _ in range(batch):
            invoke()
    out = []
    for _ in range(repetitions):
        prepare()
        total = timer(invoke, batch)
        out.append((total, total / batch))
    return out


def run_scenario(ax, torch, timer, cfg, degree, bits, emit):
    n = degree
    q = params.find_ntt_prime(degree, bits)
    root = params.find_primitive_2n_root(degree, q)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()

    result = torch.empty(n, dtype=torch.int64, device="cuda")
    none = lambda: None  # noqa: E731
    sel = set(cfg.kernels)

    def rng_for(name):
        return MT19937_64(scenario_seed(cfg.seed, degree, bits, name))

    def run(name, invoke, prepare=none):
        batch = cfg.batch if not name.startswith("ntt") else 1
        samples = measure(timer, cfg.warmup, cfg.repetitions, batch, prepare, invoke)
        emit(name, degree, bits, q, samples, batch)

    for name, fn in (("add_vv", ax.EltwiseAddMod), ("sub_vv", ax.EltwiseSubMod)):
        if name in sel:
            g = rng_for(name)
            a, b = dev(g.factor_values(n, q, 1)), dev(g.factor_values(n, q, 1))
            run(name, lambda fn=fn, a=a, b=b: fn(result, a, b, n, q))
    for name, fn in (("add_vs", ax.EltwiseAddMod), ("sub_vs", ax.EltwiseSubMod)):
        if name in sel:
            g = rng_for(name)
            a = dev(g.factor_values(n, q, 1))
            s = int(g.bounded(1, q)[0])
            run(name, lambda fn=fn, a=a, s=s: fn(result, a, s, n, q))
    for f in (1, 2, 4):
        name = "mul_f%d" % f
        if name in sel:
            g = rng_for(name)
            a, b = dev(g.factor_values(n, q, f)), dev(g.factor_values(n, q, f))
            run(name, lambda a=a, b=b, f=f: ax.EltwiseMulMod(result, a, b, n, q, f))
    for f in (1, 2, 4, 8):
        name = "fma_f%d" % f
        if name in sel:
            g = rng_for(name)
            a, c = dev(g.factor_values(n, q, f)), dev(g.factor_values(n, q, f))
            s = int(g.factor_values(1, q, f)[0])
            run(name, lambda a=a, c=c, s=s, f=f: ax.EltwiseFMAMod(result, a, s, c, n, q, f))
    for name, fin, fout in (("reduce_2to1", 2, 1), ("reduce_4to1", 4, 1), ("reduce_4to2", 4, 2),
                            ("reduce_barrett", None, 1)):
        if name in sel:
            g = rng_for(name)
            a = dev(g.draws(n).copy() if fin is None else g.factor_values(n, q, fin))
            pin = q if fin is None else fin
            run(name, lambda a=a, pin=pin, fout=fout: ax.EltwiseReduceMod(result, a, n, q, pin, fout))
    ntt_names = [k for k in ("ntt_forward", "ntt_forward_lazy", "ntt_inverse", "ntt_inverse_lazy") if k in sel]
    if ntt_names:
        plan = ax.NTT(n, q, root)  # table generation outside every timed interval
        base = dev(rng_for("ntt_base").factor_values(n, q, 1))
        inv_base = base.clone()
        plan.ComputeForward(inv_base, inv_base, 1, 1)
        work = torch.empty_like(base)
        for name, fwd, fo in (("ntt_forward", True, 1), ("ntt_forward_lazy", True, 4),
                              ("ntt_inverse", False, 1), ("ntt_inverse_lazy", False, 2)):
            if name not in sel:
                continue
            src = base if fwd else inv_base
            call = plan.ComputeForward if fwd else plan.ComputeInverse
            run(name, lambda call=call, fo=fo: call(work, work, 1, fo), prepare=lambda src=src: work.copy_(src))


class Writers:
    def __init__(self, out_dir, write_raw, batch, timer_mode):
        os.makedirs(out_dir, exist_ok=True)
        self.timer_mode = timer_mode
        self.sf = open(os.path.join(out_dir, "summary.csv"), "w", newline="")
        self.summary = csv.writer(self.sf, lineterminator="\n")
        self.summary.writerow(SUMMARY_COLUMNS)
        self.rf = None
        if write_raw:
            self.rf = open(os.path.join(out_dir, "raw_samples.csv"), "w", newline="")
            self.raw = csv.writer(self.rf, lineterminator="\n")
            self.raw.writerow(RAW_COLUMNS)

    def emit(self, name, degree, bits, q, samples, batch):
        mf, fin, fout = kernel_meta(name, q)
        key = [BACKEND, name, degree, bits, q, mf, fin, fout]
        if self.rf:
            for i, (total, per) in enumerate(samples):
                self.raw.writerow(key + [i, batch, "none", "%.3f" % total, "%.3f" % total, "%.3f" % total,
                                         "%.3f" % per, 0, 0, 0, "0.000", 0, 0, 0, 0, 0, 0, 0])
        st = statistics([per for _, per in samples])
        f = lambda x: "%.3f" % x  # noqa: E731
        self.summary.writerow(key + [st["samples_total"], st["samples_kept"], st["outliers_removed"],
                                     f(st["raw_min"]), f(st["raw_max"]), f(st["q1"]), f(st["q3"]), f(st["iqr"]),
                                     f(st["lower_fence"]), f(st["upper_fence"]), f(st["min"]), f(st["max"]),
                                     f(st["mean"]), f(st["median"]), f(st["stddev"]), f(st["cv_percent"]),
                                     f(st["mean"] / degree), f(st["median"] / degree), 0])
        self.sf.flush()
        print("  %-20s median=%.1f ns  mean=%.1f  kept=%d/%d  CV=%.2f%%" % (
            name, st["median"], st["mean"], st["samples_kept"], st["samples_total"], st["cv_percent"]), flush=True)

    def close(self):
        self.sf.close()
        if self.rf:
            self.rf.close()


def write_metadata(cfg, out_dir, torch):
    lines = [
        "backend=" + BACKEND,
        "device=" + torch.cuda.get_device_name(0),
        "torch=" + torch.__version__,
        "host=" + platform.node(),
        "date=" + datetime.datetime.now().isoformat(timespec="seconds"),
        "repetitions=%d" % cfg.repetitions,
        "warmup=%d" % cfg.warmup,
        "batch=%d" % cfg.batch,
        "seed=0x%X" % cfg.seed,
        "degrees=" + ",".join(map(str, cfg.degrees)),
        "qi_bits=" + ",".join(map(str, cfg.bits)),
        "kernels=" + ",".join(cfg.kernels),
        "pmu_profile=none",
        "counter_unit=nanoseconds per invocation (the *_cycles columns hold ns, not CPU cycles)",
        "timer=" + ("host: perf_counter_ns around the call(s) and a stream synchronisation" if cfg.timer == "host"
                    else "device: HIP events on the call's stream around the call(s)"),
        "operands=device-resident, generated on the host by mt19937_64 per scenario and kernel",
        "outlier_filter=Tukey IQR, fences Q1-1.5*IQR and Q3+1.5*IQR",
        "quantiles=linear interpolation over sorted samples",
        "ntt_mode=in-place; NTT table construction excluded from measurement",
        "batch_cycles_semantics=raw_cycles is the total of a batch sample; cycles/statistics are per invocation",
    ]
    with open(os.path.join(out_dir, "metadata.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


def compare(gpu_summary, cpu_summary, cpu_ghz, output):
    """Merge a GPU summary with a reference CPU summary: CPU cycles become ns
    at cpu_ghz, and each row reports CPU median / GPU median."""
    keys = ["kernel", "poly_degree", "qi_bits", "modulus", "mod_factor", "input_mod_factor", "output_mod_factor"]

    def load(p):
        with open(p, newline="") as fh:
            return {tuple(r[k] for k in keys): r for r in csv.DictReader(fh)}

    g, c = load(gpu_summary), load(cpu_summary)
    common = sorted(set(g) & set(c), key=lambda k: (k[0], int(k[1]), int(k[2])))
    fields = keys + ["gpu_backend", "gpu_median_ns", "gpu_mean_ns", "cpu_backend", "cpu_median_cycles",
                     "cpu_median_ns", "speedup_median_cpu_over_gpu", "speedup_mean_cpu_over_gpu"]
    with open(output, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields, lineterminator="\n")
        w.writeheader()
        for k in common:
            gr, cr = g[k], c[k]
            cmed = float(cr["median_cycles"]) / cpu_ghz
            cmean = float(cr["mean_cycles"]) / cpu_ghz
            gmed, gmean = float(gr["median_cycles"]), float(gr["mean_cycles"])
            row = dict(zip(keys, k))
            row.update({"gpu_backend": gr["backend"], "gpu_median_ns": gr["median_cycles"],
                        "gpu_mean_ns": gr["mean_cycles"], "cpu_backend": cr["backend"],
                        "cpu_median_cycles": cr["median_cycles"], "cpu_median_ns": "%.3f" % cmed,
                        "speedup_median_cpu_over_gpu": "%.6f" % (cmed / gmed) if gmed else "",
                        "speedup_mean_cpu_over_gpu": "%.6f" % (cmean / gmean) if gmean else ""})
            w.writerow(row)
    return len(common)


def parse_args(argv):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--repetitions", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch", type=int, default=1, help="invocations per sample (element-wise kernels only)")
    ap.add_argument("--degrees", default=",".join(map(str, DEFAULT_DEGREES)))
    ap.add_argument("--bits", default=",".join(map(str, DEFAULT_BITS)))
    ap.add_argument("--kernels", default="all")
    ap.add_argument("--seed", type=lambda s: int(s, 0), default=DEFAULT_SEED)
    ap.add_argument("--shuffle-scenarios", action="store_true")
    ap.add_argument("--shuffle-seed", type=lambda s: int(s, 0), default=0)
    ap.add_argument("--timer", choices=["device", "host"], default="device")
    ap.add_argument("--output-dir", default=None)
    ap.add_argument("--no-raw", action="store_true")
    ap.add_argument("--list-kernels", action="store_true")
    ap.add_argument("--compare", nargs=2, metavar=("GPU_SUMMARY", "CPU_SUMMARY"))
    ap.add_argument("--cpu-ghz", type=float, default=None)
    ap.add_argument("--output", default=None, help="--compare output CSV")
    cfg = ap.parse_args(argv)
    cfg.degrees = [int(x) for x in cfg.degrees.split(",") if x.strip()]
    cfg.bits = [int(x) for x in cfg.bits.split(",") if x.strip()]
    cfg.kernels = KERNELS[:] if cfg.kernels == "all" else [k.strip() for k in cfg.kernels.split(",") if k.strip()]
    # the reference's validation (axhel_benchmark.cpp:566-585)
    if cfg.repetitions < 4:
        ap.error("--repetitions must be >= 4 for IQR statistics")
    if cfg.batch < 1:
        ap.error("--batch must be >= 1")
    bad = [k for k in cfg.kernels if k not in KERNELS]
    if bad:
        ap.error("unknown kernels: " + ",".join(bad))
    if not cfg.degrees or not cfg.bits or not cfg.kernels:
        ap.error("empty benchmark matrix")
    for d in cfg.degrees:
        if d < 2 or d & (d - 1):
            ap.error("all polynomial degrees must be powers of two >= 2")
    for b in cfg.bits:
        if not 3 <= b <= 62:
            ap.error("modulus bit widths must be in [3,62]")
    if cfg.compare and (cfg.cpu_ghz is None or cfg.output is None):
        ap.error("--compare needs --cpu-ghz and --output")
    return cfg


def scenarios(cfg):
    """(degree, bits) pairs in the reference's fixed order; --shuffle-scenarios
    permutes them with Python's generator (deterministic per seed, not the
    reference's std::shuffle order)."""
    out = [(d, b) for d in cfg.degrees for b in cfg.bits]
    if cfg.shuffle_scenarios:
        random.Random(cfg.shuffle_seed).shuffle(out)
    return out


def main(argv=None):
    cfg = parse_args(sys.argv[1:] if argv is None else argv)
    if cfg.list_kernels:
        print("\n".join(KERNELS))
        return 0
    if cfg.compare:
        rows = compare(cfg.compare[0], cfg.compare[1], cfg.cpu_ghz, cfg.output)
        print("Wrote %s (%d scenarios)" % (cfg.output, rows))
        return 0
    import torch
    import axhel_amd as ax
    if not torch.cuda.is_available():
        raise SystemExit("the scenario benchmark needs a GPU")
    out_dir = cfg.output_dir or os.path.join("benchmark_results", datetime.datetime.now().strftime("%Y%m%d-%H%M%S"),
                                             "gpu")
    os.makedirs(out_dir, exist_ok=True)
    write_metadata(cfg, out_dir, torch)
    w = Writers(out_dir, not cfg.no_raw, cfg.batch, cfg.timer)
    timer = Timer(torch, cfg.timer)
    sc = scenarios(cfg)
    print("AXHEL scenario benchmark backend: %s  repetitions %d, warm-up %d, batch %d, timer %s\nOutput: %s" % (
        BACKEND, cfg.repetitions, cfg.warmup, cfg.batch, cfg.timer, out_dir), flush=True)
    try:
        for i, (d, b) in enumerate(sc):
            print("\nScenario %d/%d N=%d, qi_bits=%d, q=%d" % (i + 1, len(sc), d, b, params.find_ntt_prime(d, b)),
                  flush=True)
            run_scenario(ax, torch, timer, cfg, d, b, w.emit)
    finally:
        w.close()
    print("\nCompleted.")
    return 0


if __name__ == "__main__":
    sys.exit(main())
