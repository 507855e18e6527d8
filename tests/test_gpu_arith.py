"""White-box tests of the device arithmetic primitives (axhel_amd/csrc/arith.hpp)
through the test hooks of libaxhel_arith_hooks.so: every inline-assembly form
(carry-out multiply-add, carry-selected conditional subtraction, borrow
chain, shift-add) and every butterfly the kernels use equals exact
integer arithmetic word for word, on edge words and random words.  The
range and congruence checks follow the reference's own primitive tests
(tests/internal/test_internal_native.cpp:26-158): they hold by mathematics,
independently of how the formulas are written."""
import ctypes
import os

import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from tests.helpers import MODULUS_BITS, boundary_values

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
(MULHI64, SHOUP_ACC, CSUB_N, SUB64, TWICE_PLUS, FWD, FWD_SKIP, FWD_SUB8, FWD_LAST, INV, INV_FINAL, BARRETT_MUL,
 BARRETT_REDUCE) = range(13)

_hooks = None


def hooks():
    global _hooks
    if _hooks is None:
        ax.lib()  # loads torch's HIP runtime first, as for the product library
        lib = ctypes.CDLL(os.path.join(ax.LIB_DIR, "libaxhel_arith_hooks.so"))
        lib.axhel_arith_hooks_apply.restype = ctypes.c_int
        lib.axhel_arith_hooks_apply.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        _hooks = lib
    return _hooks


def apply(op, items, barrett=(0, 0, 0)):
    """items: list of up-to-7-tuples of Python ints; returns a list of (out0, out1)."""
    arr = np.zeros((len(items), 7), dtype=np.uint64)
    for i, t in enumerate(items):
        arr[i, :len(t)] = [x & M64 for x in t]
    out = np.zeros((len(items), 2), dtype=np.uint64)
    assert hooks().axhel_arith_hooks_apply(op, arr.ctypes.data, out.ctypes.data, len(items), *barrett) == 0
    return [(int(a), int(b)) for a, b in out]


def shoup(y, w, wq, q):
    return (y * w - ((y * wq) >> 64) * q) & M64


def cs(x, b):
    return x - b if x >= b else x


def edge_words(q):
    v = {0, 1, 2, q - 1, q, q + 1, 2 * q - 1, 2 * q, 4 * q - 1, 4 * q, 8 * q - 1, (1 << 32) - 1, 1 << 32,
         (1 << 63) - 1, 1 << 63, M64}
    return sorted(x & M64 for x in v)


def moduli():
    return [params.find_ntt_prime(1024, b) for b in MODULUS_BITS]


def randoms(rng, count, bound):
    return [int(x) for x in params.uniform(rng, count, bound)]


def test_word_primitives_equal_exact_arithmetic():
    rng = np.random.default_rng(1)
    for q in moduli():
        e = edge_words(q)
        r = randoms(rng, 512, 1 << 64)
        pairs = [(a, b) for a in e for b in e] + list(zip(r[:256], r[256:]))
        got = apply(MULHI64, pairs)
        assert [g[0] for g in got] == [(a * b) >> 64 for a, b in pairs], q
        got = apply(SUB64, pairs)
        assert [g[0] for g in got] == [(a - b) & M64 for a, b in pairs], q
        got = apply(TWICE_PLUS, pairs)
        assert [g[0] for g in got] == [(2 * a + b) & M64 for a, b in pairs], q
        nz = [(a, b) for a, b in pairs if b != 0]  # csub_n takes nb = 2^64 - b, b > 0
        got = apply(CSUB_N, nz)
        assert [g[0] for g in got] == [cs(a, b) for a, b in nz], q


def test_shoup_product_equals_exact_arithmetic_and_is_lazy():
    """shoup_acc(acc, y, w, w', q) == acc + y w - floor(y w' / 2^64) q mod 2^64
    for any words; for a valid pair (w < q, w' = floor(w 2^64 / q)) the
    product is below 2q and congruent to y w (test_internal_native.cpp:83-124)."""
    rng = np.random.default_rng(2)
    for q in moduli():
        ws = [0, 1, 2, q // 2, q - 2, q - 1] + randoms(rng, 4, q)
        ys = boundary_values(q, 4) + edge_words(q) + randoms(rng, 64, 4 * q) + randoms(rng, 16, 1 << 64)
        items = []
        for w in ws:
            wq = (w << 64) // q
            for y in ys:
                items.append((0, y, w, wq, q))
        got = apply(SHOUP_ACC, items)
        for (acc, y, w, wq, _), (g, _) in zip(items, got):
            assert g == shoup(y, w, wq, q), (q, y, w)
            assert g < 2 * q and g % q == y * w % q, (q, y, w)
        # accumulator and arbitrary quotient words: the exact formula still holds
        r = randoms(rng, 5 * 256, 1 << 64)
        items = [(r[i], r[i + 1], r[i + 2], r[i + 3], q) for i in range(0, len(r), 5)]
        got = apply(SHOUP_ACC, items)
        assert [g[0] for g in got] == [(a + shoup(y, w, wq, q)) & M64 for a, y, w, wq, _ in items], q


def fwd_ref(x, y, w, wq, q, pre):
    xh = pre(x, q)
    t = shoup(y, w, wq, q)
    return (xh + t) & M64, (xh + 2 * q - t) & M64


PRE = {FWD: lambda x, q: cs(x, 2 * q), FWD_SKIP: lambda x, q: x, FWD_SUB8: lambda x, q: cs(x, 8 * q),
       FWD_LAST: lambda x, q: cs(cs(cs(x, 8 * q), 4 * q), 2 * q)}


def test_forward_butterflies_equal_the_harvey_formula():
    """fwd_bfly and the three relaxed stage kinds equal the reference's
    forward butterfly (ntt-internal.hpp:20-31; relaxed: the subtraction
    before it varied) on every word, valid twiddle pairs and arbitrary ones."""
    rng = np.random.default_rng(3)
    for q in moduli():
        psi = params.find_primitive_2n_root(1024, q)
        w = [psi, 1, q - 1] + randoms(rng, 2, q)
        e = edge_words(q)
        items = [(x, y, wi, (wi << 64) // q, q) for wi in w[:2] for x in e for y in e]
        r = randoms(rng, 4 * 256, 1 << 64)
        items += [(r[i], r[i + 1], r[i + 2], r[i + 3], q) for i in range(0, len(r), 4)]
        for op in (FWD, FWD_SKIP, FWD_SUB8, FWD_LAST):
            if op != FWD and q >= 1 << 60:
                continue  # the relaxed schedule runs only for moduli below 2^60
            got = apply(op, items)
            want = [fwd_ref(x, y, wi, wq, q, PRE[op]) for x, y, wi, wq, _ in items]
            assert got == want, (q, op)


def test_forward_butterfly_range_and_congruence():
    """test_internal_native.cpp:127-146: inputs below 4q give outputs below
    4q with x' = x + y W and y' = x - y W (mod q)."""
    for q in moduli():
        psi = params.find_primitive_2n_root(1024, q)
        wq = (psi << 64) // q
        vals = boundary_values(q, 4)
        items = [(x, y, psi, wq, q) for x in vals for y in vals]
        for (x, y, *_), (a, b) in zip(items, apply(FWD, items)):
            t = y * psi % q
            assert a < 4 * q and b < 4 * q, (q, x, y)
            assert a % q == (x + t) % q and b % q == (x - t) % q, (q, x, y)
        if q < 1 << 60:
            # relaxed stages: SKIP takes x < 14q to below 16q; SUB8 and LAST
            # take x < 16q to below 10q and 4q (y: any word below 16q)
            vals16 = boundary_values(q, 16)
            items = [(x, y, psi, wq, q) for x in vals16 for y in vals16]
            for op, bound in ((FWD_SUB8, 10 * q), (FWD_LAST, 4 * q)):
                for (x, y, *_), (a, b) in zip(items, apply(op, items)):
                    t = y * psi % q
                    assert a < bound and b < bound, (q, op, x, y)
                    assert a % q == (x + t) % q and b % q == (x - t) % q, (q, op, x, y)
            items = [(x, y, psi, wq, q) for x in vals16 if x < 14 * q for y in vals16]
            for op, bound in ((FWD_SKIP, 16 * q),):
                for (x, y, *_), (a, b) in zip(items, apply(op, items)):
                    t = y * psi % q
                    assert a < bound and b < bound, (q, op, x, y)
                    assert a % q == (x + t) % q and b % q == (x - t) % q, (q, op, x, y)


def test_inverse_butterflies_equal_formula_range_and_congruence():
    """inv_bfly and the scaled final stage equal the reference's formulas
    (ntt-internal.hpp:37-49, ntt-native.cpp:126-141) on every word; inputs
    below 2q give outputs below 2q with the Gentleman-Sande congruences
    (test_internal_native.cpp:148-158)."""
    rng = np.random.default_rng(4)
    for q in moduli():
        psi = params.find_primitive_2n_root(1024, q)
        wq = (psi << 64) // q
        e = edge_words(q)
        r = randoms(rng, 4 * 256, 1 << 64)
        items = [(x, y, psi, wq, q) for x in e for y in e] + [(r[i], r[i + 1], r[i + 2], r[i + 3], q)
                                                            for i in range(0, len(r), 4)]
        got = apply(INV, items)
        want = [(cs((x + y) & M64, 2 * q), shoup((x + 2 * q - y) & M64, w, w2, q)) for x, y, w, w2, _ in items]
        assert got == want, q
        vals = boundary_values(q, 2)
        items = [(x, y, psi, wq, q) for x in vals for y in vals]
        for (x, y, *_), (a, b) in zip(items, apply(INV, items)):
            assert a < 2 * q and b < 2 * q, (q, x, y)
            assert a % q == (x + y) % q and b % q == (x - y) * psi % q, (q, x, y)
        # final stage: n^-1 and the scaled root s with their quotients
        n = 1024
        ni = pow(n, q - 2, q)
        s = (ni * pow(psi, q - 2, q)) % q
        fin = [(x, y, ni, (ni << 64) // q, q, s, (s << 64) // q) for x in vals for y in vals]
        fin += [(r[i], r[i + 1], r[i + 2], r[i + 3], q, r[i] ^ r[i + 3], r[i + 1] ^ r[i + 2])
                for i in range(0, 512, 4)]
        got = apply(INV_FINAL, fin)
        for (x, y, n1, n1q, _, s1, s1q), (a, b) in zip(fin, got):
            xh = cs(x, 2 * q)
            assert (a, b) == (shoup((xh + y) & M64, n1, n1q, q), shoup((xh + 2 * q - y) & M64, s1, s1q, q)), q
        for (x, y, *_), (a, b) in zip(fin[:len(vals) ** 2], got):
            assert a < 2 * q and b < 2 * q
            assert a % q == (x + y) * ni % q and b % q == (x - y) * s % q, (q, x, y)


def barrett_consts(q):
    n = q.bit_length()
    s = n - 2
    return ((1 << (s + 64)) // q) & M64, s, (64 - s) & 63


def test_barrett_primitives():
    """barrett_mul and barrett_reduce64 equal eltwise-mul-mod.cpp:34-61 and
    eltwise-reduce-mod.cpp:19-29 word for word, and give x y mod q (operands
    below q) and x mod q (any 64-bit x)."""
    rng = np.random.default_rng(5)
    for q in moduli():
        mu, s, sc = barrett_consts(q)
        e = [x for x in edge_words(q) if x < q] + randoms(rng, 64, q)
        pairs = [(a, b, 0, 0, q) for a in e for b in e]
        got = apply(BARRETT_MUL, pairs, (mu, s, sc))
        for (a, b, *_), (g, _) in zip(pairs, got):
            assert g == a * b % q, (q, a, b)
        xs = edge_words(q) + randoms(rng, 512, 1 << 64)
        got = apply(BARRETT_REDUCE, [(x, 0, 0, 0, q) for x in xs], (mu, s, sc))
        for x, (g, _) in zip(xs, got):
            assert g == x % q, (q, x)
