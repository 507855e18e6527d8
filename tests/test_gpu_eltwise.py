"""Element-wise kernels on the GPU, bit-exact against the oracle
(tests/public/test_eltwise.cpp and test_reduce.cpp restated)."""
import numpy as np
import pytest

import axhel_amd as ax
from oracle import oracle as O
from tests.helpers import (MODULUS_BITS, boundary_values, dev, element_sizes, empty_dev, host, make_pattern)
from axhel_amd import params

pytestmark = pytest.mark.gpu


def _q(bits):
    return params.find_ntt_prime(1024, bits)


@pytest.mark.parametrize("bits", MODULUS_BITS)
def test_add_sub_boundary_all_sizes(bits):
    q = _q(bits)
    vals = boundary_values(q, 1)
    for n in element_sizes():
        a = make_pattern(n, vals)
        b = make_pattern(n, vals[::-1])
        da, db, dr = dev(a), dev(b), empty_dev(n)
        ax.EltwiseAddMod(dr, da, db, n, q)
        assert np.array_equal(host(dr), O.add_mod(a, b, q)), ("add", n)
        ax.EltwiseSubMod(dr, da, db, n, q)
        assert np.array_equal(host(dr), O.sub_mod(a, b, q)), ("sub", n)
        s = vals[n % len(vals)]
        ax.EltwiseAddMod(dr, da, s, n, q)
        assert np.array_equal(host(dr), O.add_mod(a, s, q)), ("add-vs", n)
        ax.EltwiseSubMod(dr, da, s, n, q)
        assert np.array_equal(host(dr), O.sub_mod(a, s, q)), ("sub-vs", n)


@pytest.mark.parametrize("q", [2, 3, 5, 17, 257])
def test_exhaustive_small_moduli(q):
    a = np.array([x for x in range(q) for _ in range(q)], dtype=np.uint64)
    b = np.array([y for _ in range(q) for y in range(q)], dtype=np.uint64)
    n = a.size
    da, db, dr = dev(a), dev(b), empty_dev(n)
    ax.EltwiseAddMod(dr, da, db, n, q)
    assert np.array_equal(host(dr), O.add_mod(a, b, q))
    ax.EltwiseSubMod(dr, da, db, n, q)
    assert np.array_equal(host(dr), O.sub_mod(a, b, q))
    ax.EltwiseMulMod(dr, da, db, n, q, 1)
    assert np.array_equal(host(dr), O.mul_mod(a, b, q, 1))
    ax.EltwiseFMAMod(dr, da, 1 % q, db, n, q, 1)
    assert np.array_equal(host(dr), O.fma_mod(a, 1 % q, b, q, 1))


@pytest.mark.parametrize("bits", list(range(20, 63)))
def test_mul_fma_every_barrett_width(bits):
    q = _q(bits)
    rng = np.random.default_rng(bits)
    n = 4099
    for f in (1, 2, 4):
        a, b = params.uniform(rng, n, f * q), params.uniform(rng, n, f * q)
        da, db, dr = dev(a), dev(b), empty_dev(n)
        assert ax.EltwiseMulMod(dr, da, db, n, q, f)
        assert np.array_equal(host(dr), O.mul_mod(a, b, q, f)), f
    for f in (1, 2, 4, 8):
        bound = min(f * q, 1 << 64)
        a, c = params.uniform(rng, n, bound), params.uniform(rng, n, bound)
        s = int(params.uniform(rng, 1, bound)[0])
        da, dc, dr = dev(a), dev(c), empty_dev(n)
        ax.EltwiseFMAMod(dr, da, s, dc, n, q, f)
        assert np.array_equal(host(dr), O.fma_mod(a, s, c, q, f)), ("fma", f)
        ax.EltwiseFMAMod(dr, da, s, None, n, q, f)
        assert np.array_equal(host(dr), O.fma_mod(a, s, None, q, f)), ("fma-null", f)


@pytest.mark.parametrize("bits", MODULUS_BITS)
def test_mul_boundaries_all_sizes(bits):
    q = _q(bits)
    for f in (1, 2, 4):
        vals = boundary_values(q, f)
        for n in element_sizes():
            a, b = make_pattern(n, vals), make_pattern(n, vals[::-1])
            da, db, dr = dev(a), dev(b), empty_dev(n)
            ax.EltwiseMulMod(dr, da, db, n, q, f)
            assert np.array_equal(host(dr), O.mul_mod(a, b, q, f)), (f, n)


@pytest.mark.parametrize("bits", MODULUS_BITS)
def test_reduce_all_modes(bits):
    q = _q(bits)
    rng = np.random.default_rng(bits)
    for fin, fout in ((1, 1), (1, 2), (2, 2), (2, 1), (4, 1), (4, 2)):
        vals = boundary_values(q, fin)
        for n in (0, 1, 7, 33, 1025, 8193):
            a = np.concatenate([make_pattern(n // 2, vals), params.uniform(rng, n - n // 2, fin * q)])
            da, dr = dev(a), empty_dev(n)
            ax.EltwiseReduceMod(dr, da, n, q, fin, fout)
            assert np.array_equal(host(dr), O.reduce_mod(a, q, fin, fout)), (fin, fout, n)
    # Barrett mode, the whole 64-bit domain (SEAL's EltwiseReduceMod(..., q, 1))
    M = (1 << 64) - 1
    edge = np.array([x for x in (0, 1, q - 1, q, q + 1, 2 * q, 4 * q - 1, 7 * q + 1, M, M - 1, M - q) if x <= M],
                    dtype=np.uint64)
    a = np.concatenate([edge, params.uniform(rng, 100003, 1 << 64)])
    da, dr = dev(a), empty_dev(a.size)
    for fout in (1, 2):
        ax.EltwiseReduceMod(dr, da, a.size, q, q, fout)
        assert np.array_equal(host(dr), O.reduce_mod(a, q, q, fout))


def test_aliasing_in_place():
    """test_eltwise.cpp:439-516: result may alias an operand exactly."""
    q = _q(50)
    rng = np.random.default_rng(5)
    n = 10001
    a, b = params.uniform(rng, n, q), params.uniform(rng, n, q)
    for op in ("add", "sub", "mul", "fma", "red"):
        da, db = dev(a), dev(b)
        if op == "add":
            ax.EltwiseAddMod(da, da, db, n, q)
            want = O.add_mod(a, b, q)
        elif op == "sub":
            ax.EltwiseSubMod(db, da, db, n, q)
            da = db
            want = O.sub_mod(a, b, q)
        elif op == "mul":
            ax.EltwiseMulMod(da, da, da, n, q, 1)
            want = O.mul_mod(a, a, q, 1)
        elif op == "fma":
            ax.EltwiseFMAMod(db, da, 7, db, n, q, 1)
            da = db
            want = O.fma_mod(a, 7, b, q, 1)
        else:
            ax.EltwiseReduceMod(da, da, n, q, q, 1)
            want = O.reduce_mod(a, q, q, 1)
        assert np.array_equal(host(da), want), op


def test_unaligned_offsets_and_canaries():
    """test_eltwise.cpp:518-635: 8-byte aligned views and untouched guards."""
    torch = pytest.importorskip("torch")
    q = _q(60)
    rng = np.random.default_rng(6)
    guard = 8
    for n in (1, 2, 3, 31, 32, 33, 1025, 8193):
        a = params.uniform(rng, n, q)
        b = params.uniform(rng, n, q)
        for off in (0, 1):
            big_a = dev(np.concatenate([np.zeros(off, np.uint64), a]))
            big_b = dev(np.concatenate([np.zeros(1 - off, np.uint64), b]))
            out = dev(np.full(n + 2 * guard + 1, 0x13579BDF2468ACE0, dtype=np.uint64))
            va, vb = big_a[off:], big_b[1 - off:]
            vr = out[guard + off: guard + off + n]
            ax.EltwiseMulMod(vr, va, vb, n, q, 1)
            got = host(out)
            assert np.array_equal(got[guard + off: guard + off + n], O.mul_mod(a, b, q, 1))
            assert (got[: guard + off] == 0x13579BDF2468ACE0).all()
            assert (got[guard + off + n:] == 0x13579BDF2468ACE0).all()


def test_host_pointers_are_staged():
    q = _q(52)
    rng = np.random.default_rng(8)
    n = 5000
    a, b = params.uniform(rng, n, 4 * q), params.uniform(rng, n, 4 * q)
    r = np.zeros(n, dtype=np.uint64)
    ax.EltwiseMulMod(r, a, b, n, q, 4)
    assert np.array_equal(r, O.mul_mod(a, b, q, 4))
    c = a.copy()
    ax.EltwiseReduceMod(c, c, n, q, 4, 1)  # in place on a host buffer
    assert np.array_equal(c, O.reduce_mod(a, q, 4, 1))


def test_unsupported_factors_leave_output_untouched():
    q = _q(50)
    n = 64
    a = dev(np.arange(n, dtype=np.uint64))
    sentinel = np.full(n, 4242, dtype=np.uint64)
    for f in (0, 3, 8):
        r = dev(sentinel)
        assert ax.EltwiseMulMod(r, a, a, n, q, f) is False
        assert np.array_equal(host(r), sentinel)
    for f in (0, 3, 16):
        r = dev(sentinel)
        assert ax.EltwiseFMAMod(r, a, 3, a, n, q, f) is False
        assert np.array_equal(host(r), sentinel)
    for fin, fout in ((2, 4), (8, 1)):
        r = dev(sentinel)
        assert ax.EltwiseReduceMod(r, a, n, q, fin, fout) is False
        assert np.array_equal(host(r), sentinel)


def test_reference_kats_on_gpu():
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
    for c in g["mul"]:
        r = empty_dev(8)
        ax.EltwiseMulMod(r, dev(c["a"]), dev(c["b"]), 8, c["q"], c["f"])
        assert host(r).tolist() == c["out"]
    for c in g["fma"]:
        r = empty_dev(8)
        ax.EltwiseFMAMod(r, dev(c["a"]), c["s"], None if c["c"] is None else dev(c["c"]), 8, c["q"], c["f"])
        assert host(r).tolist() == c["out"]
    for c in g["reduce"]:
        n = len(c["a"])
        r = empty_dev(n)
        ax.EltwiseReduceMod(r, dev(c["a"]), n, c["q"], c["in_f"], c["out_f"])
        assert host(r).tolist() == c["out"]


def test_large_stream_checksum():
    """Full-size run of each op at 2^26 elements against the oracle."""
    q = _q(59)
    rng = np.random.default_rng(9)
    n = 1 << 26
    a, b = params.uniform(rng, n, q), params.uniform(rng, n, q)
    da, db, dr = dev(a), dev(b), empty_dev(n)
    ax.EltwiseMulMod(dr, da, db, n, q, 1)
    assert np.array_equal(host(dr), O.mul_mod(a, b, q, 1))
    ax.EltwiseFMAMod(dr, da, 12345, db, n, q, 1)
    assert np.array_equal(host(dr), O.fma_mod(a, 12345, b, q, 1))
    ax.EltwiseAddMod(dr, da, db, n, q)
    assert np.array_equal(host(dr), O.add_mod(a, b, q))
