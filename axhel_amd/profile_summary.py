"""Text summaries of rocprofv3 runs of bench.py, as kept in profiles/.

    python -m axhel_amd.profile_summary trace    --dir D --bench JSON --command CMD
    python -m axhel_amd.profile_summary hbm      --fetch D --write D --copy-fetch D --copy-write D --command CMD
    python -m axhel_amd.profile_summary counters --dir D --bench JSON --command CMD

trace reads a --kernel-trace --stats run (run_kernel_stats.csv,
run_kernel_trace.csv) and the bench line the same run printed; hbm reads
--pmc FETCH_SIZE and --pmc WRITE_SIZE runs of the bench and of
`bench.py --calibrate-copy` (one EltwiseReduceMod 4->2 over 2^27 words:
1 GiB read, 1 GiB written), which scales each counter to bytes; counters
reads a --pmc run of SQ_/GRBM_ counters.  The two headline pass kernels are
named by --pass0 / --pass1 (defaults: the 64 x 24 x 2^16 forward batch's COL
and stacked ROW passes on the relaxed schedule).  The last lines of the hbm
and counters summaries (`traffic_bytes ...`, `valu_insts ...`) are what
bench.py reads back for the workload they cover.
"""
import argparse
import csv
import json
import os
import statistics
import sys
from collections import defaultdict

PASS0 = "axhel_dev::ntt_pass_kernel<8, 4, 16, true, true, false, true>"
PASS1 = "axhel_dev::ntt_row_stack_kernel<8, 4, 16, true, true>"


def short(name):
    """Kernel name without 'void ', anonymous-namespace qualifiers and its
    parameter list."""
    n = name[5:] if name.startswith("void ") else name
    n = n.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def bench_line(path):
    with open(path) as f:
        return json.loads(f.read().strip().splitlines()[-1])


def counters_by_kernel(d):
    """{(kernel, grid): {counter: mean over dispatches}} from run_counter_collection.csv."""
    acc = defaultdict(lambda: defaultdict(list))
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            acc[(short(r["Kernel_Name"]), int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: statistics.fmean(v) for c, v in cs.items()} for k, cs in acc.items()}


def header(title, command, extra=()):
    out = [f"# {title}", f"# command: {command}"]
    out += [f"# {x}" for x in extra]
    return out


def cmd_trace(a):
    b = bench_line(a.bench)
    rows = []
    with open(os.path.join(a.dir, "run_kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            rows.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                         float(r["MinNs"]), float(r["MaxNs"])))
    rows.sort(key=lambda x: -x[2])
    out = header(a.title, a.command, ["columns from run_kernel_stats.csv (average over every dispatch of the run, "
                                      "warm-up and variants included)"])
    out.append(f"{'kernel':66s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s}")
    for name, calls, _, avg, mn, mx in [r for r in rows if r[0]][:a.top]:
        out.append(f"{name[:66]:66s} {calls:6d} {avg / 1e3:9.2f} {mn / 1e3:9.2f} {mx / 1e3:9.2f}")
    # per-dispatch durations of the headline passes; the bench's instrumented
    # copy of its timed steps is the last `steps` dispatches of each
    durs = defaultdict(list)
    with open(os.path.join(a.dir, "run_kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            durs[short(r["Kernel_Name"])].append((int(r["Start_Timestamp"]),
                                                  int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    steps = int(b["steps"])
    out.append("#")
    kernels = b["kernels"]
    for p, name in enumerate((a.pass0, a.pass1)):
        d = [x[1] for x in sorted(durs.get(name, []))]
        if not d:
            out.append(f"# pass {p}: {name} not in the trace")
            continue
        inst = statistics.fmean(d[-steps:]) / 1e3
        allv = statistics.fmean(d) / 1e3
        k = kernels[p]
        bfly, byts, peak = k["info"]["butterflies"], k["info"]["bytes"], k["valu"]["peak"]
        out.append(f"# headline pass {p} ({k['info']['schedule_name']}, stacked={k['info']['stacked']}): {name}")
        out.append(f"#   all {len(d)} dispatches: avg {allv:.2f} us; the bench's instrumented copy of the timed steps "
                   f"(last {steps} dispatches): avg {inst:.2f} us")
        out.append(f"#   recomputed: {bfly} butterflies / {inst:.2f} us = {bfly / inst / 1e3:.1f} Gbfly/s = "
                   f"{bfly / inst / 1e3 / peak:.4f} of the {peak:.1f} Gbfly/s schedule peak; {byts} B / {inst:.2f} us "
                   f"= {byts / inst / 1e3:.0f} GB/s = {byts / inst / 1e3 / 8000:.4f} of 8 TB/s")
    r = b["roofline"]
    out.append(f"# bench line of the same run: kernel ms {[round(x, 4) for x in b['kernel_ms']]}, "
               f"roofline bound {r['bound']}, frac {r['frac']:.4f} (valu {r['valu']['frac']:.4f}, "
               f"hbm {r['hbm']['frac']:.4f}); headline step {b['ms_per_step']:.4f} ms ({b['value']:.4g} limbs/s)")
    print("\n".join(out))


def cmd_hbm(a):
    fetch, write = counters_by_kernel(a.fetch), counters_by_kernel(a.write)
    cf, cw = counters_by_kernel(a.copy_fetch), counters_by_kernel(a.copy_write)
    true_kib = (1 << 27) * 8 / 1024  # one GiB each way
    ff = true_kib / max(v["FETCH_SIZE"] for v in cf.values())
    fw = true_kib / max(v["WRITE_SIZE"] for v in cw.values())
    out = header(a.title, a.command, [f"copy calibration: {a.copy_command}",
                                      "(EltwiseReduceMod 4->2 over 2^27 words: 1048576 KiB read and 1048576 KiB "
                                      "written)",
                                      "counter values in KiB, averaged over the dispatches of each (kernel, grid)"])
    out.append(f"{'kernel':66s} {'grid':>8s} {'FETCH_SIZE':>12s} {'WRITE_SIZE':>12s}")
    keys = [k for k in fetch if k in write and (k[0].startswith("axhel_dev::ntt") or "polymul" in k[0])]
    # the batch's launches (the largest grid); single-limb calls of the same run are left out
    top = max((k[1] for k in keys), default=0)
    keys = sorted(k for k in keys if k[1] == top)
    for k in keys:
        out.append(f"{k[0][:66]:66s} {k[1]:8d} {fetch[k]['FETCH_SIZE']:12.1f} {write[k]['WRITE_SIZE']:12.1f}")
    out.append("#")
    out.append(f"# scale factors from the copy: FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f}")
    alg = a.limbs * a.degree * 16 / 1024
    out.append(f"# algorithmic bytes per NTT pass launch: {a.limbs} limbs x {a.degree} words x 16 B (read + write) "
               f"= {alg:.0f} KiB")
    traffic = {}
    for k in keys:
        t = fetch[k]["FETCH_SIZE"] * ff + write[k]["WRITE_SIZE"] * fw
        out.append(f"# {k[0][:60]:60s} fetch {fetch[k]['FETCH_SIZE'] * ff:9.0f} + write "
                   f"{write[k]['WRITE_SIZE'] * fw:9.0f} = {t:9.0f} KiB = {t / alg:.3f}x algorithmic")
        for p, name in enumerate((a.pass0, a.pass1)):
            if k[0] == name:
                traffic[p] = t * 1024
    out.append("# forward step = COL pass (pass 0) + stacked ROW pass (pass 1)")
    out.append("traffic_bytes " + " ".join(f"pass{p}={traffic[p]:.0f}" for p in sorted(traffic)))
    print("\n".join(out))


def cmd_counters(a):
    b = bench_line(a.bench)
    c = counters_by_kernel(a.dir)
    out = header(a.title, a.command, ["counter values per dispatch, averaged over the dispatches of each "
                                      "(kernel, grid); SQ_ counters summed over the device, GRBM_GUI_ACTIVE over "
                                      "the 8 XCDs"])
    names = sorted({n for cs in c.values() for n in cs})
    out.append(f"{'kernel':66s} {'grid':>8s} {'counter':28s} {'avg_value':>18s}")
    for k in sorted(c):
        if not (k[0].startswith("axhel_dev::ntt") or "bfly_peak" in k[0]):
            continue
        for n in names:
            if n in c[k]:
                out.append(f"{k[0][:66]:66s} {k[1]:8d} {n:28s} {c[k][n]:18.1f}")
    out.append("#")
    out.append("# derived, per launch of the headline passes")
    insts = {}
    for p, name in enumerate((a.pass0, a.pass1)):
        ks = [k for k in c if k[0] == name]
        if not ks:
            continue
        v = c[ks[0]]
        bfly = b["kernels"][p]["info"]["butterflies"]
        per_simd_cycles = v["GRBM_GUI_ACTIVE"] / 8
        busy = v["SQ_INSTS_VALU"] * 4 / (per_simd_cycles * 1024)
        dual = v.get("SQ_ACTIVE_INST_VALU2", float("nan")) / v["SQ_INSTS_VALU"]
        lds = v.get("SQ_INSTS_LDS", float("nan"))
        out.append(f"# pass {p} {name}: {v['SQ_INSTS_VALU'] * 64 / bfly:.2f} VALU per butterfly; VALU busy {busy:.2f} "
                   f"(SQ_INSTS_VALU x 4 cycles over GRBM_GUI_ACTIVE/8 x 1024 SIMDs); second issues "
                   f"{dual:.3f} of VALU instructions (SQ_ACTIVE_INST_VALU2); LDS bank conflict cycles "
                   f"{v.get('SQ_LDS_BANK_CONFLICT', float('nan')):.0f} over {lds:.0f} LDS instructions")
        insts[p] = v["SQ_INSTS_VALU"]
    out.append("valu_insts " + " ".join(f"pass{p}={insts[p]:.0f}" for p in sorted(insts)))
    print("\n".join(out))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="what", required=True)
    for name in ("trace", "hbm", "counters"):
        s = sub.add_parser(name)
        s.add_argument("--title", required=True)
        s.add_argument("--command", required=True, help="the command that produced the data, as run")
        s.add_argument("--pass0", default=PASS0)
        s.add_argument("--pass1", default=PASS1)
        if name in ("trace", "counters"):
            s.add_argument("--dir", required=True)
            s.add_argument("--bench", required=True, help="the JSON line bench.py printed in the same run")
        if name == "trace":
            s.add_argument("--top", type=int, default=24)
        if name == "hbm":
            for x in ("fetch", "write", "copy-fetch", "copy-write"):
                s.add_argument("--" + x, required=True)
            s.add_argument("--copy-command", required=True)
            s.add_argument("--limbs", type=int, default=1536)
            s.add_argument("--degree", type=int, default=65536)
    a = ap.parse_args(argv)
    {"trace": cmd_trace, "hbm": cmd_hbm, "counters": cmd_counters}[a.what](a)


if __name__ == "__main__":
    sys.exit(main())
