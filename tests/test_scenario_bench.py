"""Scenario benchmark in the reference benchmark's report format
(benchmark/src/axhel_benchmark.cpp): operand generator, statistics, report
schema; one small GPU run."""
import csv
import os

import numpy as np
import pytest

from axhel_amd import scenario_bench as sb
from oracle import oracle as O


def test_mt19937_64_matches_the_standard_and_the_oracle():
    g = sb.MT19937_64(5489)  # default seed: the standard fixes the 10000th output
    assert int(g.draws(10000)[-1]) == 9981545732273789042
    for seed in (0, 1, sb.DEFAULT_SEED, sb.scenario_seed(sb.DEFAULT_SEED, 4096, 50, "mul_f4")):
        ours = sb.MT19937_64(seed)
        ref = O.MT19937_64(seed)
        assert [ours.next() for _ in range(700)] == [ref.next() for _ in range(700)]
        # bounded draws consume the same stream positions (rejections included)
        for bound in (17, (1 << 63) + 12345, 3 * ((1 << 61) - 1), 4 * 1152921504606830593):
            assert np.array_equal(ours.bounded(1000, bound), ref.random_vector(1000, bound)), (seed, bound)
        assert ours.next() == ref.next()


def test_factor_values_take_raw_draws_when_the_bound_reaches_2_64():
    q = 4611686018427322369  # 62-bit: 8q > 2^64
    a, b = sb.MT19937_64(7), sb.MT19937_64(7)
    assert np.array_equal(a.factor_values(100, q, 8), b.draws(100))
    assert bool((sb.MT19937_64(9).factor_values(1000, q, 2) < 2 * q).all())


def test_scenario_seeds_are_distinct_per_kernel_and_scenario():
    seeds = {sb.scenario_seed(sb.DEFAULT_SEED, d, b, k) for d in sb.DEFAULT_DEGREES for b in sb.DEFAULT_BITS
             for k in sb.KERNELS + ["ntt_base"]}
    assert len(seeds) == len(sb.DEFAULT_DEGREES) * len(sb.DEFAULT_BITS) * (len(sb.KERNELS) + 1)
    assert all(0 <= s < 1 << 64 for s in seeds)


def test_statistics_follow_tukey_iqr():
    rng = np.random.default_rng(0)
    base = list(rng.normal(1000.0, 10.0, 97))
    data = base + [5000.0, 6000.0, 10.0]  # three outliers
    st = sb.statistics(data)
    v = np.sort(np.array(data))
    q1, q3 = np.percentile(v, 25), np.percentile(v, 75)  # numpy's default is the same linear rule
    assert st["q1"] == pytest.approx(q1) and st["q3"] == pytest.approx(q3)
    kept = v[(v >= q1 - 1.5 * (q3 - q1)) & (v <= q3 + 1.5 * (q3 - q1))]
    assert st["samples_total"] == 100 and st["samples_kept"] == kept.size == 97 and st["outliers_removed"] == 3
    assert st["mean"] == pytest.approx(kept.mean()) and st["median"] == pytest.approx(np.median(kept))
    assert st["stddev"] == pytest.approx(kept.std(ddof=1))
    assert st["cv_percent"] == pytest.approx(100 * kept.std(ddof=1) / kept.mean())
    one = sb.statistics([3.0])
    assert one["median"] == 3.0 and one["stddev"] == 0.0


def test_argument_validation_mirrors_the_reference():
    assert sb.parse_args([]).kernels == sb.KERNELS
    for bad in (["--repetitions", "3"], ["--degrees", "1000"], ["--bits", "63"], ["--kernels", "nope"],
                ["--batch", "0"]):
        with pytest.raises(SystemExit):
            sb.parse_args(bad)


def test_compare_converts_cpu_cycles(tmp_path):
    row = {c: "0" for c in sb.SUMMARY_COLUMNS}
    row.update(kernel="ntt_forward", poly_degree="4096", qi_bits="50", modulus="97")
    g = dict(row, backend=sb.BACKEND, median_cycles="1000.0", mean_cycles="1100.0")
    c = dict(row, backend="Native", median_cycles="30000.0", mean_cycles="33000.0")
    for name, r in (("g.csv", g), ("c.csv", c)):
        with open(tmp_path / name, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=sb.SUMMARY_COLUMNS)
            w.writeheader()
            w.writerow(r)
    assert sb.compare(tmp_path / "g.csv", tmp_path / "c.csv", 3.0, tmp_path / "out.csv") == 1
    out = list(csv.DictReader(open(tmp_path / "out.csv")))[0]
    assert float(out["cpu_median_ns"]) == 10000.0 and float(out["speedup_median_cpu_over_gpu"]) == 10.0


@pytest.mark.gpu
def test_small_gpu_run_writes_the_reference_schema(tmp_path):
    pytest.importorskip("torch")
    out = str(tmp_path / "gpu")
    assert sb.main(["--degrees", "2048", "--bits", "30,62", "--repetitions", "6", "--warmup", "2",
                    "--output-dir", out]) == 0
    rows = list(csv.DictReader(open(os.path.join(out, "summary.csv"))))
    assert list(rows[0].keys()) == sb.SUMMARY_COLUMNS
    assert len(rows) == 2 * len(sb.KERNELS)
    assert {r["kernel"] for r in rows} == set(sb.KERNELS)
    for r in rows:
        assert r["backend"] == sb.BACKEND and int(r["samples_total"]) == 6
        assert 0 < int(r["samples_kept"]) <= 6 and float(r["median_cycles"]) > 0
    raw = list(csv.DictReader(open(os.path.join(out, "raw_samples.csv"))))
    assert list(raw[0].keys()) == sb.RAW_COLUMNS and len(raw) == 6 * len(rows)
    meta = open(os.path.join(out, "metadata.txt")).read()
    assert "counter_unit=nanoseconds" in meta


@pytest.mark.gpu
def test_host_timer_batch_subset_and_shuffle(tmp_path):
    pytest.importorskip("torch")
    out = str(tmp_path / "gpu2")
    kernels = ["add_vs", "fma_f8", "reduce_barrett", "ntt_inverse_lazy"]
    assert sb.main(["--degrees", "4096,2048", "--bits", "40,61", "--repetitions", "5", "--warmup", "1",
                    "--batch", "3", "--timer", "host", "--kernels", ",".join(kernels), "--shuffle-scenarios",
                    "--shuffle-seed", "7", "--no-raw", "--output-dir", out]) == 0
    rows = list(csv.DictReader(open(os.path.join(out, "summary.csv"))))
    assert len(rows) == 4 * len(kernels)
    assert {(r["poly_degree"], r["qi_bits"]) for r in rows} == {(d, b) for d in ("4096", "2048") for b in ("40", "61")}
    assert all(r["kernel"] in kernels and float(r["median_cycles"]) > 0 for r in rows)
    assert not os.path.exists(os.path.join(out, "raw_samples.csv"))
    meta = open(os.path.join(out, "metadata.txt")).read()
    assert "timer=host" in meta and "batch=3" in meta
