"""Batched RNS polynomial product on the GPU (fused and as the plain call
sequence), bit-exact against the oracle running the reference's calls:
ComputeForward(a), ComputeForward(b), EltwiseMulMod(F=1), EltwiseFMAMod(F=1),
ComputeInverse (tests/public/test_polymul.cpp:13-25 plus the FMA step)."""
import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import dev, empty_dev, host, torch_mod

pytestmark = pytest.mark.gpu


def oracle_polymul(a, b, c, s, t, q):
    A = O.compute_forward(a, t, 1, 1)
    B = O.compute_forward(b, t, 1, 1)
    C = O.mul_mod(A, B, q, 1)
    D = O.fma_mod(C, s, c, q, 1)
    return O.compute_inverse(D, t, 1, 1)


def run_case(n, limbs, fused, with_addend, seed):
    torch = torch_mod()
    rng = np.random.default_rng(seed)
    qs = params.rns_primes(n, limbs)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [ax.NTT(n, q, p) for q, p in zip(qs, psis)]
    A = np.stack([params.uniform(rng, n, q) for q in qs])
    B = np.stack([params.uniform(rng, n, q) for q in qs])
    C = np.stack([params.uniform(rng, n, q) for q in qs])
    S = [int(params.uniform(rng, 1, q)[0]) for q in qs]
    da, db, dc = dev(A.ravel()).view(limbs, n), dev(B.ravel()).view(limbs, n), dev(C.ravel()).view(limbs, n)
    out = empty_dev(limbs * n).view(limbs, n)
    scr = empty_dev(limbs * n).view(limbs, n)
    jobs = [(plans[l], out[l], da[l], db[l], dc[l] if with_addend else None, S[l], scr[l]) for l in range(limbs)]
    pm = ax.Polymul(jobs, fused=fused)
    pm.launch()
    got = host(out).reshape(limbs, n)
    for l in range(limbs):
        t = O.Tables(n, qs[l], psis[l])
        want = oracle_polymul(A[l], B[l], C[l] if with_addend else None, S[l], t, qs[l])
        assert np.array_equal(got[l], want), (n, l, fused, with_addend)
    # inputs untouched
    assert np.array_equal(host(da).reshape(limbs, n), A)
    assert np.array_equal(host(db).reshape(limbs, n), B)
    pm.close()


@pytest.mark.parametrize("logn", [10, 12, 13, 14, 15, 16, 17, 18, 20])
@pytest.mark.parametrize("fused", [True, False])
def test_polymul_matches_call_sequence(logn, fused):
    run_case(1 << logn, 3, fused, True, logn)


@pytest.mark.parametrize("fused", [True, False])
def test_polymul_null_addend(fused):
    run_case(1 << 15, 2, fused, False, 7)


def test_polymul_every_config_n32768():
    n = 1 << 15
    try:
        for ci in range(ax.config_count(n)):
            ax.config_select(n, ci)
            run_case(n, 2, True, True, 100 + ci)
    finally:
        ax.config_select(n, -1)


def test_polymul_is_the_negacyclic_product():
    """scalar 1, no addend: the result is a*b mod (X^N + 1)."""
    torch = torch_mod()
    n = 1 << 12
    q = params.find_ntt_prime(n, 60)
    psi = params.find_primitive_2n_root(n, q)
    ntt = ax.NTT(n, q, psi)
    rng = np.random.default_rng(3)
    a = np.zeros(n, dtype=np.uint64)
    b = params.uniform(rng, n, q)
    a[0], a[1], a[n - 1] = 5, 1, q - 2  # 5 + x - 2 x^(N-1)
    out, scr = empty_dev(n), empty_dev(n)
    pm = ax.Polymul([(ntt, out, dev(a), dev(b), None, 1, scr)], fused=True)
    pm.launch()
    got = host(out)
    bi = [int(x) for x in b]
    want = []
    for k in range(n):
        v = 5 * bi[k]
        v += bi[k - 1] if k >= 1 else -bi[n - 1]          # x * b: shift with negacyclic wrap
        v += 2 * bi[k + 1] if k + 1 < n else -2 * bi[0]  # -2 x^(N-1) * b: x^(N-1+j) = -x^(j-1)
        want.append(v % q)
    assert got.tolist() == want


@pytest.mark.parametrize("logn", [15, 16])
def test_polymul_stacked_every_config(logn):
    """Many products per prime: the fused kernel runs over stacks of same-plan
    jobs (one LDS twiddle set per workgroup, a padded last stack per prime).
    Every job equals the five-call sequence; one job per prime equals the
    oracle."""
    torch = torch_mod()
    n = 1 << logn
    qs = params.rns_primes(n, 2)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [ax.NTT(n, q, p) for q, p in zip(qs, psis)]
    rng = np.random.default_rng(logn + 50)
    try:
        for ci in range(ax.config_count(n)):
            ax.config_select(n, ci)
            gpw = ax.config_get(n, ci)["row_groups"]
            counts = [4 * gpw, 3 * gpw + 1]
            owner = [pi for pi, c in enumerate(counts) for _ in range(c)]
            J = len(owner)
            A = np.stack([params.uniform(rng, n, qs[pi]) for pi in owner])
            B = np.stack([params.uniform(rng, n, qs[pi]) for pi in owner])
            C = np.stack([params.uniform(rng, n, qs[pi]) for pi in owner])
            S = [int(params.uniform(rng, 1, qs[pi])[0]) for pi in owner]
            da, db, dc = (dev(X.ravel()).view(J, n) for X in (A, B, C))
            outs, scrs = [], []
            for fused in (True, False):
                out = empty_dev(J * n).view(J, n)
                scr = empty_dev(J * n).view(J, n)
                pm = ax.Polymul([(plans[pi], out[i], da[i], db[i], dc[i], S[i], scr[i])
                                 for i, pi in enumerate(owner)], fused=fused)
                if fused:
                    assert pm.stacked_groups() == 1, (logn, ci)
                pm.launch()
                torch.cuda.synchronize()
                pm.close()
                outs.append(out)
                scrs.append(scr)
            assert torch.equal(outs[0], outs[1]), (logn, ci)
            got = host(outs[0]).reshape(J, n)
            for pi in range(len(counts)):
                i = owner.index(pi) + counts[pi] - 1  # last job of the prime: in the padded stack
                t = O.Tables(n, qs[pi], psis[pi])
                assert np.array_equal(got[i], oracle_polymul(A[i], B[i], C[i], S[i], t, qs[pi])), (logn, ci, pi)
            assert np.array_equal(host(da).reshape(J, n), A)
    finally:
        ax.config_select(n, -1)


def test_random_mixed_polymul_jobs():
    """Seeded random product jobs in one call: degrees 2^10..2^17 (single- and
    two-pass fused kernels), moduli of 30..62 bits (both sides of the relaxed
    schedule's 2^60 gate), with and without an addend, every buffer only
    8-byte aligned.  Fused equals the five-call sequence on every job; a
    sample equals the oracle."""
    torch = torch_mod()
    rng = np.random.default_rng(77)
    specs, total = [], 0
    for _ in range(40):
        logn = int(rng.integers(10, 18))
        specs.append((1 << logn, int(rng.integers(30, 63)), bool(rng.integers(0, 2))))
        total += 6 * (1 << logn) + 1
    plans = {}
    for n, bits, _ in specs:
        if (n, bits) not in plans:
            q = params.find_ntt_prime(n, bits)
            psi = params.find_primitive_2n_root(n, q)
            plans[(n, bits)] = (ax.NTT(n, q, psi), q, psi)
    base = torch.zeros(total + 1, dtype=torch.int64, device="cuda")
    arena = base[1:]
    outs = {}
    placed = []
    off = 0
    for n, bits, has_c in specs:
        plan, q, psi = plans[(n, bits)]
        a, b, c = (params.uniform(rng, n, q) for _ in range(3))
        views = [arena[off + k * n:off + (k + 1) * n] for k in range(6)]  # a, b, c, fused out, seq out, scratch
        for v, x in zip(views[:3], (a, b, c)):
            v.copy_(torch.from_numpy(x.view(np.int64)))
        s = int(params.uniform(rng, 1, q)[0])
        placed.append((plan, q, psi, n, has_c, s, a, b, c, views))
        off += 6 * n + 1
    for fused, slot in ((True, 3), (False, 4)):
        pm = ax.Polymul([(p[0], p[9][slot], p[9][0], p[9][1], p[9][2] if p[4] else None, p[5], p[9][5])
                         for p in placed], fused=fused)
        pm.launch()
        torch.cuda.synchronize()
        pm.close()
    for i, (plan, q, psi, n, has_c, s, a, b, c, views) in enumerate(placed):
        assert torch.equal(views[3], views[4]), (i, n, q.bit_length(), has_c)
    for i in range(0, len(placed), 6):
        plan, q, psi, n, has_c, s, a, b, c, views = placed[i]
        t = O.Tables(n, q, psi)
        A, B = O.compute_forward(a, t, 1, 1), O.compute_forward(b, t, 1, 1)
        D = O.fma_mod(O.mul_mod(A, B, q, 1), s, c if has_c else None, q, 1)
        assert np.array_equal(host(views[3]), O.compute_inverse(D, t, 1, 1)), (i, n, q.bit_length(), has_c)


@pytest.mark.parametrize("create_cfg,launch_cfg", [(1, 2), (2, 1)])
def test_polymul_keeps_the_configuration_it_was_created_with(create_cfg, launch_cfg):
    """At 2^13 configurations 0-1 run one pass and 2-3 run two, and the fused
    product's descriptors depend on the pass count.  A product created under
    one and launched after the degree's selection changed to the other still
    equals the five-call sequence."""
    torch = torch_mod()
    n, limbs = 1 << 13, 3
    qs = params.rns_primes(n, limbs)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [ax.NTT(n, q, p) for q, p in zip(qs, psis)]
    rng = np.random.default_rng(41)
    A, B, C = (np.stack([params.uniform(rng, n, q) for q in qs]) for _ in range(3))
    S = [int(params.uniform(rng, 1, q)[0]) for q in qs]
    da, db, dc = (dev(X.ravel()).view(limbs, n) for X in (A, B, C))
    out = empty_dev(limbs * n).view(limbs, n)
    scr = empty_dev(limbs * n).view(limbs, n)
    jobs = [(plans[l], out[l], da[l], db[l], dc[l], S[l], scr[l]) for l in range(limbs)]
    try:
        ax.config_select(n, create_cfg)
        pm = ax.Polymul(jobs, fused=True)
        ax.config_select(n, launch_cfg)
        pm.launch()
        got = host(out).reshape(limbs, n)
        pm.close()
    finally:
        ax.config_select(n, -1)
    for l in range(limbs):
        want = oracle_polymul(A[l], B[l], C[l], S[l], O.Tables(n, qs[l], psis[l]), qs[l])
        assert np.array_equal(got[l], want), (create_cfg, launch_cfg, l)
