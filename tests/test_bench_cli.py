"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message before anything touches a device, and the serving loop's
fixed-load p99 figures are lifted beside the headline unchanged."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        out = subprocess.run([sys.executable, "bench.py", "--gpus", n], cwd=ROOT, capture_output=True,
                             text=True, timeout=120)
        assert out.returncode == 2, out.stdout + out.stderr
        assert "one GPU only" in out.stdout + out.stderr
        assert '"metric"' not in out.stdout  # no bench line for an unsupported run


def test_p99_summary_lifts_every_fixed_load():
    serving = {"mix": "m", "loads": [
        {"fraction": 0.5, "offered_req_per_s": 100.0, "p50_us": 1.0, "p99_us": 2.0, "seconds": 4},
        {"fraction": 0.8, "offered_req_per_s": 160.0, "p50_us": 3.0, "p99_us": 9.0, "seconds": 4}]}
    got = bench.p99_summary(serving)
    assert got == [
        {"offered_req_per_s": 100.0, "fraction_of_capacity": 0.5, "p50_us": 1.0, "p99_us": 2.0, "mix": "m"},
        {"offered_req_per_s": 160.0, "fraction_of_capacity": 0.8, "p50_us": 3.0, "p99_us": 9.0, "mix": "m"}]
    assert bench.p99_summary(None) is None and bench.p99_summary({}) is None
