"""axhel_amd.profile_summary on small synthetic rocprofv3 outputs (CPU):
the lines bench.py reads back (traffic_bytes, valu_insts) come out of the
counters as stated, and kernel names are shortened the same way everywhere."""
import csv
import json
import os

from axhel_amd import profile_summary as PS

P0 = "void axhel_dev::ntt_pass_kernel<8, 4, 16, true, true, false, true>(axhel_dev::PassArgs)"
P1 = "void axhel_dev::ntt_row_stack_kernel<8, 4, 16, true, true>(axhel_dev::PassArgs)"


def write_counters(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)


def bench_json(path):
    k = {"info": {"butterflies": 402653184, "bytes": 1610612736, "schedule_name": "x", "stacked": 1},
         "valu": {"peak": 2000.0}}
    line = {"steps": 2, "kernels": [k, k], "kernel_ms": [0.28, 0.32], "ms_per_step": 0.6, "value": 2.5e6,
            "roofline": {"bound": "hbm", "frac": 0.62, "valu": {"frac": 0.61}, "hbm": {"frac": 0.62}}}
    with open(path, "w") as f:
        f.write("log line\n" + json.dumps(line) + "\n")


def test_short_names():
    assert PS.short(P0) == "axhel_dev::ntt_pass_kernel<8, 4, 16, true, true, false, true>"
    assert PS.short("void (anonymous namespace)::bfly_peak_kernel<2>(unsigned long*)") == "bfly_peak_kernel<2>"


def test_hbm_traffic_line(tmp_path, capsys):
    g = 6291456
    d = str(tmp_path)
    write_counters(d + "/f", [(P0, g, "FETCH_SIZE", 393216.0), (P1, g, "FETCH_SIZE", 405504.0),
                              (P0, 8192, "FETCH_SIZE", 300.0)])
    write_counters(d + "/w", [(P0, g, "WRITE_SIZE", 786432.0), (P1, g, "WRITE_SIZE", 786432.0),
                              (P0, 8192, "WRITE_SIZE", 512.0)])
    write_counters(d + "/cf", [("void elt(int)", 1, "FETCH_SIZE", 524288.0)])  # half of 1 GiB: factor 2
    write_counters(d + "/cw", [("void elt(int)", 1, "WRITE_SIZE", 1048576.0)])
    PS.main(["hbm", "--title", "t", "--command", "c", "--copy-command", "cc", "--fetch", d + "/f", "--write",
             d + "/w", "--copy-fetch", d + "/cf", "--copy-write", d + "/cw"])
    out = capsys.readouterr().out.splitlines()
    assert "FETCH_SIZE x 2.000, WRITE_SIZE x 1.000" in "\n".join(out)
    assert out[-1] == f"traffic_bytes pass0={(393216 * 2 + 786432) * 1024} pass1={(405504 * 2 + 786432) * 1024}"
    assert not any(" 8192 " in x for x in out)  # the single-limb launches are left out


def test_counters_valu_line(tmp_path, capsys):
    d = str(tmp_path)
    rows = []
    for k, insts in ((P0, 110493696.0), (P1, 137625600.0)):
        rows += [(k, 6291456, "SQ_INSTS_VALU", insts), (k, 6291456, "GRBM_GUI_ACTIVE", 4.0e6),
                 (k, 6291456, "SQ_ACTIVE_INST_VALU2", insts / 16), (k, 6291456, "SQ_INSTS_LDS", 1.0),
                 (k, 6291456, "SQ_LDS_BANK_CONFLICT", 0.0)]
    write_counters(d + "/c", rows)
    bench_json(d + "/b.json")
    PS.main(["counters", "--title", "t", "--command", "c", "--dir", d + "/c", "--bench", d + "/b.json"])
    out = capsys.readouterr().out.splitlines()
    assert out[-1] == "valu_insts pass0=110493696 pass1=137625600"
    assert "17.56 VALU per butterfly" in out[-3] and "21.88 VALU per butterfly" in out[-2]
