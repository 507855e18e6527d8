"""Inputs outside the reference's stated domains: element-wise operands at or
above q (or F*q), NTT inputs at or above 4q (forward) or 2q (inverse), up to
full 64-bit words.  The reference checks none of them; its formulas then wrap
modulo 2^64 (eltwise-add-mod.cpp:15-47, eltwise-sub-mod.cpp:15-47,
eltwise-mul-mod.cpp:19-63, eltwise-fma-mod.cpp:19-78,
eltwise-reduce-mod.cpp:32-128, ntt-native.cpp:19-142).  The GPU kernels
evaluate the same formulas in the same order, so they equal the oracle bit
for bit on such words too.

One schedule is exempt: the normalized forward transform for moduli below
2^60 assumes the stated domain [0, 4q) (ntt-native.cpp:17) and is tested at
that domain's boundary in test_gpu_ntt.py.  Here the normalized forward runs
only with 62-bit moduli, where the kernels keep the literal formulas."""
import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import SEED, dev, empty_dev, host, ntt_params, torch_mod

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _words(rng, n, q):
    """n words over the whole 64-bit range, led by the edges of every domain."""
    edge = [M64, M64 - 1, 1 << 63, (1 << 63) - 1, 8 * q, 8 * q - 1, 4 * q, 4 * q - 1, 2 * q, 2 * q - 1,
            q, q - 1, 0, 1]
    edge = np.array([e & M64 for e in edge], dtype=np.uint64)
    return np.concatenate([edge, params.uniform(rng, n - edge.size, 1 << 64)])


@pytest.mark.parametrize("bits", [20, 40, 50, 60, 62])
def test_eltwise_out_of_contract_words(bits):
    q = params.find_ntt_prime(1024, bits)
    rng = np.random.default_rng(SEED + bits)
    for n in (4099, (1 << 16) + 3):
        a, b = _words(rng, n, q), _words(rng, n, q)[::-1].copy()
        da, db, dr = dev(a), dev(b), empty_dev(n)
        ax.EltwiseAddMod(dr, da, db, n, q)
        assert np.array_equal(host(dr), O.add_mod(a, b, q)), ("add", n)
        ax.EltwiseSubMod(dr, da, db, n, q)
        assert np.array_equal(host(dr), O.sub_mod(a, b, q)), ("sub", n)
        for s in (M64, 2 * q + 3, q):
            ax.EltwiseAddMod(dr, da, s, n, q)
            assert np.array_equal(host(dr), O.add_mod(a, s, q)), ("add-vs", n, s)
            ax.EltwiseSubMod(dr, da, s, n, q)
            assert np.array_equal(host(dr), O.sub_mod(a, s, q)), ("sub-vs", n, s)
        for f in (1, 2, 4):
            assert ax.EltwiseMulMod(dr, da, db, n, q, f)
            assert np.array_equal(host(dr), O.mul_mod(a, b, q, f)), ("mul", n, f)
        for f in (1, 2, 4, 8):
            for s in (M64, (8 * q + 5) & M64):
                ax.EltwiseFMAMod(dr, da, s, db, n, q, f)
                assert np.array_equal(host(dr), O.fma_mod(a, s, b, q, f)), ("fma", n, f, s)
                ax.EltwiseFMAMod(dr, da, s, None, n, q, f)
                assert np.array_equal(host(dr), O.fma_mod(a, s, None, q, f)), ("fma-null", n, f, s)
        for fi, fo in ((2, 1), (4, 1), (4, 2)):
            ax.EltwiseReduceMod(dr, da, n, q, fi, fo)
            assert np.array_equal(host(dr), O.reduce_mod(a, q, fi, fo)), ("reduce", n, fi, fo)


def _check_ntt(ntt, t, n, q, rng, label):
    a = _words(rng, n, q) if n >= 16 else params.uniform(rng, n, 1 << 64)
    d, r = dev(a), empty_dev(n)
    ntt.ComputeForward(r, d, 4, 4)
    assert np.array_equal(host(r), O.compute_forward(a, t, 4, 4)), ("fwd lazy", label)
    ntt.ComputeInverse(r, d, 2, 2)
    assert np.array_equal(host(r), O.compute_inverse(a, t, 2, 2)), ("inv lazy", label)
    ntt.ComputeInverse(r, d, 2, 1)
    assert np.array_equal(host(r), O.compute_inverse(a, t, 2, 1)), ("inv norm", label)
    if q >= 1 << 60:  # literal formulas for the normalized forward as well
        ntt.ComputeForward(r, d, 4, 1)
        assert np.array_equal(host(r), O.compute_forward(a, t, 4, 1)), ("fwd norm", label)


@pytest.mark.parametrize("logn", [1, 2, 4, 8, 10, 12, 13, 14, 15, 16, 17, 18, 19, 20])
def test_ntt_out_of_contract_words_every_config(logn):
    n = 1 << logn
    rng = np.random.default_rng(SEED ^ logn)
    try:
        for bits in (50, 62):
            q, psi = ntt_params(n, bits)
            ntt = ax.NTT(n, q, psi)
            t = O.Tables(n, q, psi)
            for ci in range(ax.config_count(n)):
                ax.config_select(n, ci)
                _check_ntt(ntt, t, n, q, rng, (n, bits, ci))
    finally:
        ax.config_select(n, -1)


def test_batched_lazy_ntt_out_of_contract_words():
    """The stacked second pass of a batch (several jobs per prime) keeps the
    literal formulas too: 3 polynomials x 4 limbs at N=2^16, lazy forward and
    lazy inverse, every job against the oracle."""
    torch = torch_mod()
    n, L, P = 1 << 16, 4, 3
    qs = params.rns_primes(n, L)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [ax.NTT(n, q, p) for q, p in zip(qs, psis)]
    tabs = [O.Tables(n, q, p) for q, p in zip(qs, psis)]
    rng = np.random.default_rng(SEED + 7)
    data = np.stack([np.stack([_words(rng, n, q) for q in qs]) for _ in range(P)])
    src = torch.from_numpy(data.view(np.int64).copy()).cuda()
    for inverse, of in ((False, 4), (True, 2)):
        out = torch.full_like(src, -1)
        b = ax.Batch([(plans[l], out[p, l], src[p, l], inverse, of) for p in range(P) for l in range(L)])
        b.launch_eager()
        torch.cuda.synchronize()
        b.close()
        got = out.cpu().numpy().view(np.uint64)
        for p in range(P):
            for l in range(L):
                f = O.compute_inverse if inverse else O.compute_forward
                want = f(data[p, l], tabs[l], 2 if inverse else 4, of)
                assert np.array_equal(got[p, l], want), (inverse, p, l)
