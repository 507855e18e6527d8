"""Pins the CPU oracle (oracle/axhel_oracle.c) against the reference's golden
vectors and against what the mathematics fixes, so the GPU parity tests that
compare with it mean something.  CPU only."""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from axhel_amd import params

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
M64 = (1 << 64) - 1


def brv(i, logn):
    return int(format(i, f"0{logn}b")[::-1], 2) if logn else 0


# ------------------------------------------------------------------ tables
def test_tables_match_reference_example_golden():
    k = GOLD["ntt_n8_q17"]
    t = O.Tables(k["n"], k["q"], k["psi"])
    assert t.fwd.reshape(-1, 2).tolist() == k["root_powers"]
    assert t.inv.reshape(-1, 2).tolist() == k["inv_root_powers"]
    assert t.ninv.tolist() == k["inv_degree_modulo"]


@pytest.mark.parametrize("n,bits", [(16, 30), (256, 50), (1024, 62), (4096, 59)])
def test_tables_closed_form(n, bits):
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    logn = n.bit_length() - 1
    ipsi = pow(psi, q - 2, q)
    rnd = random.Random(n)
    for i in [0, 1, 2, n - 1] + [rnd.randrange(n) for _ in range(64)]:
        w = pow(psi, brv(i, logn), q)
        assert int(t.fwd[2 * i]) == w
        assert int(t.fwd[2 * i + 1]) == (w << 64) // q
        if i:
            wi = pow(ipsi, brv(i - 1, logn) + 1, q)
            assert int(t.inv[2 * i]) == wi
            assert int(t.inv[2 * i + 1]) == (wi << 64) // q
    assert int(t.inv[0]) == 0 and int(t.inv[1]) == 0
    assert int(t.ninv[0]) * n % q == 1


# --------------------------------------------------------------- primitives
def test_shoup_textbook_bounds():
    rnd = random.Random(1)
    for bits in (20, 40, 50, 60, 61, 62):
        q = params.find_ntt_prime(1024, bits)
        for _ in range(400):
            w = rnd.randrange(q)
            wq = (w << 64) // q
            assert O.shoup_quotient(w, q) == wq
            v = rnd.randrange(1 << 64)
            t = O.shoup_lazy(v, w, wq, q)
            assert t < 2 * q and t % q == v * w % q  # Shoup's guarantee for any 64-bit v


# --------------------------------------------------------------------- NTT
def dft_closed_form(a, q, psi):
    """out[i] = a(psi^(2 brv(i) + 1)) mod q: the bit-reversed negacyclic DFT."""
    n = len(a)
    logn = n.bit_length() - 1
    out = []
    for i in range(n):
        x = pow(psi, 2 * brv(i, logn) + 1, q)
        acc, p = 0, 1
        for c in a:
            acc = (acc + int(c) * p) % q
            p = p * x % q
        out.append(acc)
    return out


def test_ntt_kat_reference():
    k = GOLD["ntt_n8_q17"]
    t = O.Tables(k["n"], k["q"], k["psi"])
    f = O.compute_forward(k["input"], t, 1, 1)
    assert f.tolist() == k["forward_normalized"]
    assert O.compute_inverse(f, t, 1, 1).tolist() == k["input"]


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 256])
@pytest.mark.parametrize("bits", [20, 50, 62])
def test_forward_closed_form_and_lazy_range(n, bits):
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    mt = O.MT19937_64(0x415848454C ^ n ^ bits)
    a = mt.random_vector(n, 4 * q)  # forward accepts [0, 4q)
    want = dft_closed_form(a, q, psi)
    lazy = O.compute_forward(a, t, 4, 4)
    norm = O.compute_forward(a, t, 4, 1)
    assert norm.tolist() == want
    assert all(int(x) < 4 * q for x in lazy)
    assert [int(x) % q for x in lazy] == want


def fold(a, q, L, zeta):
    """Coefficients of a(x) mod (x^L - zeta)."""
    out = [0] * L
    zp = 1
    for t in range(len(a) // L):
        for k in range(L):
            out[k] = (out[k] + int(a[t * L + k]) * zp) % q
        zp = zp * zeta % q
    return out


@pytest.mark.parametrize("n,bits", [(16, 30), (64, 50), (256, 62)])
def test_forward_every_stage_closed_form(n, bits):
    """After s Cooley-Tukey stages, block i (length L = n/2^s) holds
    a(x) mod (x^L - psi^((2 brv_s(i)+1) L)): pins the twiddle index and the
    pairing of every stage, and the lazy range after each."""
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    a = O.MT19937_64(7).random_vector(n, 4 * q)
    logn = n.bit_length() - 1
    for s in range(logn + 1):
        st = O.fwd_lazy(a, t, stop_log_m=s)
        L = n >> s
        assert all(int(x) < 4 * q for x in st)
        for i in range(1 << s):
            zeta = pow(psi, (2 * brv(i, s) + 1) * L, q)
            assert [int(x) % q for x in st[i * L:(i + 1) * L]] == fold(a, q, L, zeta), (s, i)


@pytest.mark.parametrize("n,bits", [(8, 20), (64, 50), (256, 62)])
def test_inverse_every_stage_undoes_forward(n, bits):
    """k Gentleman-Sande stages on the spectrum equal 2^k times the forward
    state with k fewer stages (mod q), each value < 2q; the scaled final
    stage returns the input exactly."""
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    a = O.MT19937_64(11).random_vector(n, q)
    logn = n.bit_length() - 1
    spec = O.compute_forward(a, t, 1, 1)
    spec_lazy2 = O.reduce_mod(O.compute_forward(a, t, 1, 4), q, 4, 2)
    for k in range(logn):
        st = O.inv_lazy(spec_lazy2, t, stages=k, final=False)
        assert all(int(x) < 2 * q for x in st)
        fw = O.fwd_lazy(a, t, stop_log_m=logn - k)
        assert [int(x) % q for x in st] == [(int(y) << k) % q for y in fw], k
    full = O.inv_lazy(spec_lazy2, t)
    assert all(int(x) < 2 * q for x in full)
    assert [int(x) % q for x in full] == a.tolist()
    assert O.compute_inverse(spec, t, 1, 1).tolist() == a.tolist()


def negacyclic(a, b, q):
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            p = int(a[i]) * int(b[j]) % q
            if i + j < n:
                out[i + j] = (out[i + j] + p) % q
            else:
                out[i + j - n] = (out[i + j - n] - p) % q
    return out


# ----------------------------------------- lazy representatives (Harvey 2014)
#
# The checks above fix every lazy word modulo q and its range.  Which
# congruent value a lazy transform returns is fixed by the butterfly itself:
# D. Harvey, "Faster arithmetic for number-theoretic transforms", J. Symbolic
# Computation 60 (2014) 113-119, Algorithm 4, the butterfly the reference
# names and implements (ntt-internal.hpp:20-31), and its Gentleman-Sande
# counterpart (ntt-internal.hpp:37-49).  Restated below over Python integers
# from the paper's statement, with every twiddle and Shoup quotient computed
# here from psi (not read from the oracle's tables); the last inverse stage
# is the reference's own fused scaling (ntt-native.cpp:101-141), which no
# publication fixes.  The oracle must agree word for word, lazy values
# included, on inputs that include the boundary words 0, q, 2q - 1, 2q and
# 4q - 1.

def _shoup_quot(w, q):
    return (w << 64) // q


def _shoup_exact(y, w, wq, q):
    """Harvey's Algorithm 4 lines 2-3: T = W Y - floor(W' Y / 2^64) p, an
    integer in [0, 2p) (Harvey's Lemma 2 for Y < 2^64)."""
    t = w * y - ((wq * y) >> 64) * q
    assert 0 <= t < 2 * q
    return t


def harvey_forward(a, q, psi):
    """Negacyclic Cooley-Tukey transform, bit-reversed output, with Harvey's
    lazy butterfly: inputs and outputs in [0, 4p)."""
    x = [int(v) for v in a]
    n = len(x)
    logn = n.bit_length() - 1
    m, gap = 1, n >> 1
    while m < n:
        for i in range(m):
            w = pow(psi, brv(m + i, logn), q)  # the twiddle of block i when there are m blocks
            wq = _shoup_quot(w, q)
            for j in range(i * 2 * gap, i * 2 * gap + gap):
                X, Y = x[j], x[j + gap]
                if X >= 2 * q:  # Algorithm 4 line 1
                    X -= 2 * q
                T = _shoup_exact(Y, w, wq, q)
                x[j], x[j + gap] = X + T, X - T + 2 * q
        m, gap = m << 1, gap >> 1
    return x


def gs_inverse(a, q, psi, final=True):
    """Gentleman-Sande inverse on [0, 2p) words with the lazy butterfly
    X' = X + Y reduced below 2p, Y' = W (X - Y + 2p) by Shoup; then the
    reference's last stage, which folds n^-1 into both outputs."""
    x = [int(v) for v in a]
    n = len(x)
    logn = n.bit_length() - 1
    ipsi = pow(psi, q - 2, q)
    m, gap = n >> 1, 1
    while m > 1:
        for i in range(m):
            w = pow(ipsi, brv(n - 2 * m + i, logn) + 1, q)  # table entry n - 2m + 1 + i
            wq = _shoup_quot(w, q)
            for j in range(i * 2 * gap, i * 2 * gap + gap):
                X, Y = x[j], x[j + gap]
                S = X + Y
                if S >= 2 * q:
                    S -= 2 * q
                x[j], x[j + gap] = S, _shoup_exact(X - Y + 2 * q, w, wq, q)
        m, gap = m >> 1, gap << 1
    if not final:
        return x
    ninv = pow(n, q - 2, q)
    ninv_q = _shoup_quot(ninv, q)
    wl = pow(ipsi, brv(n - 2, logn) + 1, q)  # table entry n - 1
    s = _shoup_exact(wl, ninv, ninv_q, q)
    if s >= q:
        s -= q
    sq = _shoup_quot(s, q)
    for j in range(gap):
        X, Y = x[j], x[j + gap]
        if X >= 2 * q:
            X -= 2 * q
        x[j], x[j + gap] = _shoup_exact(X + Y, ninv, ninv_q, q), _shoup_exact(X + 2 * q - Y, s, sq, q)
    return x


def _boundary_words(q, bound, n, seed):
    rng = random.Random(seed)
    edges = [0, 1, q - 1, q, 2 * q - 1, 2 * q, 4 * q - 1]
    words = [e for e in edges if e < bound]
    return [words[i] if i < len(words) else rng.randrange(bound) for i in range(n)]


@pytest.mark.parametrize("n,bits", [(2, 20), (8, 40), (64, 50), (256, 62), (1024, 60)])
def test_forward_lazy_representative_harvey(n, bits):
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    for seed in range(3):
        a = _boundary_words(q, 4 * q, n, seed * 1000 + n)
        if seed == 1:
            a = a[::-1]  # boundary words on the butterflies' other operand
        if seed == 2:
            a = [2 * q if i % 2 == 0 else v for i, v in enumerate(a)]  # x == 2q exactly in every first-stage pair
        got = O.compute_forward(np.array(a, dtype=np.uint64), t, 4, 4)
        assert got.tolist() == harvey_forward(a, q, psi), seed


@pytest.mark.parametrize("n,bits", [(2, 20), (8, 40), (64, 50), (256, 62), (1024, 60)])
def test_inverse_lazy_representative_harvey(n, bits):
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    logn = n.bit_length() - 1
    for seed in range(3):
        a = _boundary_words(q, 2 * q, n, seed * 1000 + n)
        if seed == 1:
            a = a[::-1]
        if seed == 2:
            a = [q] * n  # X + Y == 2q exactly in every first-stage pair
        arr = np.array(a, dtype=np.uint64)
        if logn > 1:
            mid = O.inv_lazy(arr, t, stages=logn - 1, final=False)
            assert mid.tolist() == gs_inverse(a, q, psi, final=False), seed
        got = O.compute_inverse(arr, t, 2, 2)
        assert got.tolist() == gs_inverse(a, q, psi), seed


@pytest.mark.parametrize("n,bits", [(8, 40), (32, 60), (64, 62)])
def test_polymul_matches_quadratic_oracle(n, bits):
    """tests/public/test_polymul.cpp:13-25 call sequence vs O(N^2) product."""
    q = params.find_ntt_prime(n, bits)
    psi = params.find_primitive_2n_root(n, q)
    t = O.Tables(n, q, psi)
    mt = O.MT19937_64(3)
    a, b = mt.random_vector(n, q), mt.random_vector(n, q)
    A, B = O.compute_forward(a, t), O.compute_forward(b, t)
    c = O.compute_inverse(O.mul_mod(A, B, q, 1), t)
    assert c.tolist() == negacyclic(a, b, q)


# ------------------------------------------------------------ element-wise
def test_eltwise_reference_kats():
    g = GOLD
    assert O.add_mod(g["add_vv"]["a"], g["add_vv"]["b"], g["add_vv"]["q"]).tolist() == g["add_vv"]["out"]
    assert O.add_mod(g["add_vs"]["a"], g["add_vs"]["s"], g["add_vs"]["q"]).tolist() == g["add_vs"]["out"]
    assert O.sub_mod(g["sub_vv"]["a"], g["sub_vv"]["b"], g["sub_vv"]["q"]).tolist() == g["sub_vv"]["out"]
    assert O.sub_mod(g["sub_vs"]["a"], g["sub_vs"]["s"], g["sub_vs"]["q"]).tolist() == g["sub_vs"]["out"]
    for c in g["mul"]:
        assert O.mul_mod(c["a"], c["b"], c["q"], c["f"]).tolist() == c["out"]
    for c in g["fma"]:
        assert O.fma_mod(c["a"], c["s"], c["c"], c["q"], c["f"]).tolist() == c["out"]
    for c in g["reduce"]:
        assert O.reduce_mod(c["a"], c["q"], c["in_f"], c["out_f"]).tolist() == c["out"]


@pytest.mark.parametrize("q", [2, 3, 5, 17, 257])
def test_eltwise_exhaustive_small_moduli(q):
    """tests/public/test_eltwise.cpp:197-250: every pair of canonical inputs."""
    xs = np.array([x for x in range(q) for _ in range(q)], dtype=np.uint64)
    ys = np.array([y for _ in range(q) for y in range(q)], dtype=np.uint64)
    X, Y = xs.tolist(), ys.tolist()
    assert O.add_mod(xs, ys, q).tolist() == [(x + y) % q for x, y in zip(X, Y)]
    assert O.sub_mod(xs, ys, q).tolist() == [(x - y) % q for x, y in zip(X, Y)]
    assert O.mul_mod(xs, ys, q, 1).tolist() == [(x * y) % q for x, y in zip(X, Y)]
    assert O.fma_mod(xs, 1 % q, ys, q, 1).tolist() == [(x + y) % q for x, y in zip(X, Y)]
    for f in (2, 4):
        lifted = np.array([x + (i % f) * q for i, x in enumerate(X)], dtype=np.uint64)
        assert O.mul_mod(lifted, ys, q, f).tolist() == [(x * y) % q for x, y in zip(X, Y)]


@pytest.mark.parametrize("bits", list(range(20, 63)))
def test_barrett_mul_all_widths(bits):
    """Every Barrett shift 18..60 (tests/public/test_eltwise.cpp:330-357)."""
    q = params.find_ntt_prime(1024, bits)
    rnd = np.random.default_rng(bits)
    for f in (1, 2, 4):
        a = params.uniform(rnd, 256, f * q)
        b = params.uniform(rnd, 256, f * q)
        got = O.mul_mod(a, b, q, f).tolist()
        assert got == [(int(x) * int(y)) % q for x, y in zip(a, b)]
    for f in (1, 2, 4, 8):
        a = params.uniform(rnd, 256, min(f * q, 1 << 64))
        c = params.uniform(rnd, 256, min(f * q, 1 << 64))
        s = int(params.uniform(rnd, 1, min(f * q, 1 << 64))[0])
        got = O.fma_mod(a, s, c, q, f).tolist()
        assert got == [(int(x) * (s % q) + int(z)) % q for x, z in zip(a, c)]
        got = O.fma_mod(a, s, None, q, f).tolist()
        assert got == [(int(x) * (s % q)) % q for x in a]


@pytest.mark.parametrize("bits", [20, 30, 40, 50, 52, 55, 59, 60, 61, 62])
def test_reduce_closed_form(bits):
    q = params.find_ntt_prime(1024, bits)
    rnd = np.random.default_rng(bits)
    for fin, fout in ((2, 1), (4, 1), (4, 2)):
        a = params.uniform(rnd, 512, fin * q)
        got = O.reduce_mod(a, q, fin, fout).tolist()
        for x, y in zip(a.tolist(), got):
            assert y % q == x % q and y < fout * q
            if fout == 1:
                assert y == x % q
    # Barrett mode: any 64-bit word, incl. the top of the domain
    a = np.concatenate([params.uniform(rnd, 1024, 1 << 64),
                        np.array([0, 1, q - 1, q, q + 1, M64, M64 - 1, M64 - q], dtype=np.uint64)])
    assert O.reduce_mod(a, q, q, 1).tolist() == [x % q for x in a.tolist()]
    assert O.reduce_mod(a, q, q, 2).tolist() == [x % q for x in a.tolist()]


def test_unsupported_factors_are_noops():
    q = params.find_ntt_prime(1024, 50)
    a = np.arange(8, dtype=np.uint64)
    sentinel = np.full(8, 12345, dtype=np.uint64)
    for f in (0, 3, 8, 16):
        out = sentinel.copy()
        O.mul_mod(a, a, q, f, out=out)
        assert out.tolist() == sentinel.tolist()
    for f in (0, 3, 5, 16):
        out = sentinel.copy()
        O.fma_mod(a, 3, a, q, f, out=out)
        assert out.tolist() == sentinel.tolist()
    for fin, fout in ((2, 4), (8, 1), (3, 1)):
        out = sentinel.copy()
        O.reduce_mod(a, q, fin, fout, out=out)
        assert out.tolist() == sentinel.tolist()


# -------------------------------------------------------------- parameters
def test_mt19937_64_standard_check_value():
    m = O.MT19937_64(5489)
    for _ in range(9999):
        m.next()
    assert m.next() == 9981545732273789042  # C++ standard [rand.predef]


@pytest.mark.parametrize("bits", [20, 30, 50, 62])
def test_prime_search(bits):
    n = 1024
    q = O.find_ntt_prime(n, bits)
    assert q == params.find_ntt_prime(n, bits)
    assert q % (2 * n) == 1 and q.bit_length() == bits and O.is_prime(q)
    c = q + 2 * n
    while c < (1 << bits):
        assert not params.is_prime(c)
        c += 2 * n
    psi = O.find_primitive_2n_root(n, q)
    assert psi == params.find_primitive_2n_root(n, q)
    assert pow(psi, n, q) == q - 1
