"""CPU model of the relaxed forward schedule (stage_kind in
axhel_amd/csrc/ntt_kernels.hpp, fwd_bfly_relaxed in arith.hpp): the
reference's butterflies on the reference's pairs and twiddles, with the
conditional subtractions placed per pass as the kernels place them, in exact
64-bit arithmetic (numpy, wrapping).  For every compiled pass split of every
degree and moduli just below 2^60 (the gate), inputs at the top of the
reference's domain (every word 4q - 1) and uniform in [0, 4q): no word ever
reaches 16q (so nothing wraps), and the normalized output equals the oracle's.
This checks the bound argument; the GPU tests check the kernels."""
import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O

M32 = np.uint64(0xFFFFFFFF)
S32 = np.uint64(32)


def mulhi64(a, b):
    """hi64(a * b) elementwise for uint64 arrays."""
    a0, a1 = a & M32, a >> S32
    b0, b1 = b & M32, b >> S32
    t = a1 * b0 + ((a0 * b0) >> S32)
    u = a0 * b1 + (t & M32)
    return a1 * b1 + (t >> S32) + (u >> S32)


def cs(x, b):
    b = np.uint64(b)
    return np.where(x >= b, x - b, x)


def stage_kind(log_g, col, p):
    s = log_g - 1 - p
    if not col and p == 0:
        return "LAST"
    if col:
        return "SUB8" if s >= 6 and (s - 6) % 4 == 0 else "SKIP"
    return "SUB8" if s % 4 == 0 else "SKIP"


def relaxed_forward(a, t, q, log_col):
    """Normalized forward transform on the relaxed schedule; returns the
    output and the largest word seen on the way."""
    n = len(a)
    logn = n.bit_length() - 1
    fw = t.fwd.reshape(-1, 2)
    x = a.astype(np.uint64).copy()
    peak = int(x.max())
    qq = np.uint64(q)
    with np.errstate(over="ignore"):
        for b in range(logn - 1, -1, -1):
            gap = 1 << b
            m = n >> (b + 1)
            if log_col and b >= logn - log_col:
                kind = stage_kind(log_col, True, b - (logn - log_col))
            else:
                kind = stage_kind(logn - log_col, False, b)
            v = x.reshape(m, 2, gap)
            lo, hi = v[:, 0, :], v[:, 1, :]
            w = fw[m:2 * m, 0][:, None]
            wq = fw[m:2 * m, 1][:, None]
            xh = lo.copy()
            if kind != "SKIP":
                xh = cs(xh, 8 * q)
            if kind == "LAST":
                xh = cs(cs(xh, 4 * q), 2 * q)
            tt = hi * w - mulhi64(hi, wq) * qq  # Shoup, wrapping
            new_lo = xh + tt
            new_hi = xh + np.uint64(2 * q) - tt
            v[:, 0, :] = new_lo
            v[:, 1, :] = new_hi
            peak = max(peak, int(x.max()))
    return cs(cs(x, 2 * q), q), peak


def splits(n):
    out = {0} if n <= 1 << 13 else set()
    for i in range(ax.config_count(n)):
        out.add(ax.config_get(n, i)["log_col"])
    return sorted(out)


@pytest.mark.parametrize("logn", [1, 2, 5, 8, 10, 12, 13, 14, 15, 16, 17, 18, 19, 20])
def test_relaxed_schedule_bound_and_result(logn):
    n = 1 << logn
    q = params.find_ntt_prime(n, 60)  # largest NTT prime below the 2^60 gate
    assert q < 1 << 60
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    rng = np.random.default_rng(logn)
    inputs = [np.full(n, 4 * q - 1, dtype=np.uint64), params.uniform(rng, n, 4 * q)]
    for log_col in splits(n):
        for a in inputs:
            got, peak = relaxed_forward(a, t, q, log_col)
            assert peak < 16 * q, (logn, log_col, peak / q)
            assert np.array_equal(got, O.compute_forward(a, t, 4, 1)), (logn, log_col)
