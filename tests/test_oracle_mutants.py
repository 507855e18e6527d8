"""Planted mistakes in the oracle, on the CPU: each mutant below is the
oracle's C source (oracle/axhel_oracle.c) with one plausible slip (a dropped
term or reduction, a wrong constant, sign, bound, index or operand),
compiled with gcc into a temporary directory.  Every mutant must fail at
least one of the checks in tests/test_oracle.py that pin the real oracle
(the reference's golden vectors, closed forms, round trips and range
invariants); the unmutated source, compiled the same way, passes them all.

The last three mutants change only which congruent value a lazy transform
returns.  The first two of them pass every congruence and range check and
fail only the restatement of Harvey's butterflies (tests/test_oracle.py,
the *_lazy_representative_harvey checks); the third also breaks the
inverse round trip.  CPU only; nothing here runs on a GPU."""
import importlib.util
import os
import subprocess

import pytest

from tests import test_oracle as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "axhel_oracle.c")
FRONT = os.path.join(ROOT, "oracle", "oracle.py")

# (name, text in the source, replacement); each text occurs exactly once
MUTANTS = [
    ("fwd: 2q offset written as q",
     "y[j] = xh + q2 - t;", "y[j] = xh + q - t;"),
    ("fwd: conditional subtraction of 2q dropped",
     "uint64_t xh = csub(x[j], q2);                 /* [0,4q) -> [0,2q) */", "uint64_t xh = x[j];"),
    ("fwd: twiddles of a stage taken in reverse order",
     "const uint64_t w = fwd[2 * (m + i)], wq = fwd[2 * (m + i) + 1];",
     "const uint64_t w = fwd[2 * (2 * m - 1 - i)], wq = fwd[2 * (2 * m - 1 - i) + 1];"),
    ("inv: difference with operands transposed",
     "y[j] = shoup_lazy(u + q2 - v, w, wq, q);", "y[j] = shoup_lazy(v + q2 - u, w, wq, q);"),
    ("inv: conditional subtraction of the sum dropped",
     "x[j] = csub(u + v, q2);", "x[j] = u + v;"),
    ("inv: final stage scales the difference by n^-1 only",
     "y[j] = shoup_lazy(dif, s, sq, q);", "y[j] = shoup_lazy(dif, ninv[0], ninv[1], q);"),
    ("inv: twiddle index starts at 0",
     "uint64_t idx = 1, m = n >> 1, gap = 1;", "uint64_t idx = 0, m = n >> 1, gap = 1;"),
    ("tables: forward powers not bit-reversed",
     "fwd[2 * k] = seq[r];\n        fwd[2 * k + 1] = oracle_shoup_quotient(seq[r], q);",
     "fwd[2 * k] = seq[k];\n        fwd[2 * k + 1] = oracle_shoup_quotient(seq[k], q);"),
    ("tables: inverse exponent off by one",
     "uint64_t w = seq[r + 1];", "uint64_t w = seq[r];"),
    ("shoup: quotient estimate one too large",
     "return v * w - qhat * q;", "return v * w - (qhat + 1) * q;"),
    ("add: s == q left unreduced",
     "uint64_t s = a[i] + b[i];\n        s -= (uint64_t)(s >= q) * q;",
     "uint64_t s = a[i] + b[i];\n        s -= (uint64_t)(s > q) * q;"),
    ("sub: q added when x == y",
     "d += (uint64_t)(x < y) * q;", "d += (uint64_t)(x <= y) * q;"),
    ("barrett: shift n - 1 instead of n - 2",
     "*s = nb - 2;", "*s = nb - 1;"),
    ("fma: addend not reduced by the input factor",
     "v += reduce_input(c[i], q, (int)f);", "v += c[i];"),
    ("reduce input F=8: first subtraction 2q instead of 4q",
     "if (f == 8) x = csub(x, 4 * q);", "if (f == 8) x = csub(x, 2 * q);"),
    ("reduce 4->1: the 2q subtraction dropped",
     "for (i = 0; i < n; ++i) r[i] = csub(csub(a[i], 2 * q), q);",
     "for (i = 0; i < n; ++i) r[i] = csub(a[i], q);"),
    ("forward: normalized output reduced 4->2 only",
     "if (out_f != 4) oracle_reduce(res, res, n, q, 4, 1);",
     "if (out_f != 4) oracle_reduce(res, res, n, q, 4, 2);"),
    ("mul: factor 8 accepted instead of a no-op",
     "if (f != 1 && f != 2 && f != 4) return;", "if (f != 1 && f != 2 && f != 4 && f != 8) return;"),
    # representative-only slips: same residues, same ranges, other lazy words
    ("fwd (representative): 2q subtracted only above 2q",
     "uint64_t xh = csub(x[j], q2);                 /* [0,4q) -> [0,2q) */",
     "uint64_t xh = x[j] > q2 ? x[j] - q2 : x[j];"),
    ("inv (representative): sum reduced only above 2q",
     "x[j] = csub(u + v, q2);", "x[j] = u + v > q2 ? u + v - q2 : u + v;"),
    ("inv final stage (representative): difference offset 4q",
     "uint64_t dif = xh + q2 - yv;", "uint64_t dif = xh + 2 * q2 - yv;"),
]

# the pinning checks of tests/test_oracle.py, at sizes that run in well under a second
CHECKS = [
    (T.test_tables_match_reference_example_golden, {}),
    (T.test_tables_closed_form, dict(n=256, bits=50)),
    (T.test_shoup_textbook_bounds, {}),
    (T.test_ntt_kat_reference, {}),
    (T.test_forward_closed_form_and_lazy_range, dict(n=16, bits=62)),
    (T.test_forward_closed_form_and_lazy_range, dict(n=64, bits=50)),
    (T.test_forward_every_stage_closed_form, dict(n=64, bits=50)),
    (T.test_inverse_every_stage_undoes_forward, dict(n=64, bits=50)),
    (T.test_inverse_every_stage_undoes_forward, dict(n=256, bits=62)),
    (T.test_polymul_matches_quadratic_oracle, dict(n=32, bits=60)),
    (T.test_eltwise_reference_kats, {}),
    (T.test_eltwise_exhaustive_small_moduli, dict(q=17)),
    (T.test_barrett_mul_all_widths, dict(bits=50)),
    (T.test_barrett_mul_all_widths, dict(bits=62)),
    (T.test_reduce_closed_form, dict(bits=50)),
    (T.test_reduce_closed_form, dict(bits=62)),
    (T.test_unsupported_factors_are_noops, {}),
    (T.test_forward_lazy_representative_harvey, dict(n=2, bits=20)),
    (T.test_forward_lazy_representative_harvey, dict(n=64, bits=50)),
    (T.test_inverse_lazy_representative_harvey, dict(n=2, bits=20)),
    (T.test_inverse_lazy_representative_harvey, dict(n=64, bits=50)),
]


def _build(tmp, name, text):
    c = os.path.join(tmp, name + ".c")
    so = os.path.join(tmp, name + ".so")
    with open(c, "w") as f:
        f.write(text)
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-shared", "-fPIC", "-o", so, c], check=True,
                   capture_output=True)
    spec = importlib.util.spec_from_file_location("oracle_" + name, FRONT)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod._LIB_PATH, mod._lib = so, None
    return mod


def _failed_checks(monkeypatch, mod):
    monkeypatch.setattr(T, "O", mod)
    failed = []
    for fn, kw in CHECKS:
        try:
            fn(**kw)
        except AssertionError:
            failed.append(fn.__name__ + str(kw or ""))
    return failed


def test_unmutated_source_passes_every_check(tmp_path, monkeypatch):
    mod = _build(str(tmp_path), "clean", open(SRC).read())
    assert _failed_checks(monkeypatch, mod) == []


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_a_check(tmp_path, monkeypatch, name, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, "mutation site must be unique"
    mod = _build(str(tmp_path), "mutant", src.replace(old, new))
    assert _failed_checks(monkeypatch, mod), "planted mistake survived every check: " + name
