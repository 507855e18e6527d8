"""The benchmark workloads at full size, checked through size-independent
properties (the oracle is too slow for every limb at these sizes):

* 64 polys x 24 limbs x N=2^16 (1536 limb-transforms, the headline batch),
  as a batch and as RNS-context calls (which must agree word for word):
  forward then inverse is the identity on every limb; the lazy forward
  reduces to the normalized forward mod q and stays below 4q; the transform
  is linear; sampled limbs equal the oracle.
* 512 jobs, N = 2^12..2^17, 1..40 limbs each, 30..60-bit moduli, both
  directions: every limb of the batch equals the plain per-call transform,
  sampled jobs equal the oracle.
* 128 polys x 32 limbs x N=2^15 RNS products: the fused path equals the
  five-call sequence on every limb, sampled limbs equal the oracle.
"""
import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import host, torch_mod

pytestmark = pytest.mark.gpu


def test_headline_batch_full_size_properties():
    torch = torch_mod()
    n, L, P = 1 << 16, 24, 64
    primes = params.rns_primes(n, L)
    psis = [params.find_primitive_2n_root(n, q) for q in primes]
    plans = [ax.NTT(n, q, p) for q, p in zip(primes, psis)]
    qt = torch.tensor(primes, dtype=torch.int64, device="cuda").view(1, L, 1)
    g = torch.Generator(device="cuda").manual_seed(0)
    src = torch.empty((P, L, n), dtype=torch.int64, device="cuda")
    for l, q in enumerate(primes):
        src[:, l, :] = torch.randint(0, q, (P, n), generator=g, device="cuda", dtype=torch.int64)
    fwd = torch.empty_like(src)
    lazy = torch.empty_like(src)
    back = torch.empty_like(src)
    jobs = lambda out, inp, inv, of: [(plans[l], out[p, l], inp[p, l], inv, of) for p in range(P) for l in range(L)]
    bf, bl, bi = ax.Batch(jobs(fwd, src, False, 1)), ax.Batch(jobs(lazy, src, False, 4)), ax.Batch(jobs(back, fwd, True, 1))
    assert bf.stacked_groups() == 1 and bl.stacked_groups() == 1 and bi.stacked_groups() == 1
    bf.launch()
    bl.launch()
    bi.launch()
    torch.cuda.synchronize()
    assert torch.equal(back, src)                         # inverse(forward(a)) == a, every limb
    assert bool((fwd < qt).all()) and bool((fwd >= 0).all())
    assert bool((lazy < 4 * qt).all()) and bool((lazy >= 0).all())
    assert torch.equal(torch.remainder(lazy, qt), fwd)     # lazy == normalized (mod q)
    # linearity on two polynomials: NTT(a + b mod q) == NTT(a) + NTT(b) mod q
    s = torch.remainder(src[0:1] + src[1:2], qt)
    s_out = torch.empty_like(s)
    bs = ax.Batch([(plans[l], s_out[0, l], s[0, l], False, 1) for l in range(L)])
    bs.launch()
    torch.cuda.synchronize()
    assert torch.equal(s_out, torch.remainder(fwd[0:1] + fwd[1:2], qt))
    # sampled limbs against the oracle
    for p, l in ((0, 0), (31, 11), (63, 23)):
        t = O.Tables(n, primes[l], psis[l])
        a = host(src[p, l])
        assert np.array_equal(host(fwd[p, l]), O.compute_forward(a, t, 1, 1)), (p, l)
        assert np.array_equal(host(lazy[p, l]), O.compute_forward(a, t, 1, 4)), (p, l)
    for b in (bf, bl, bi, bs):
        b.close()
    # the same transforms as RNS-context calls over the [P][L][N] array
    ctx = ax.RNS(n, primes, psis)
    got = torch.full_like(src, -1)
    ctx.ntt(got, src, P, False, 1)
    torch.cuda.synchronize()
    assert torch.equal(got, fwd)
    ctx.ntt(got, src, P, False, 4)
    torch.cuda.synchronize()
    assert torch.equal(got, lazy)
    ctx.ntt(got, fwd, P, True, 1)
    torch.cuda.synchronize()
    assert torch.equal(got, src)
    ctx.close()


def test_uneven_512_jobs_full_size_equal_plain_calls():
    torch = torch_mod()
    rng = np.random.default_rng(0)
    pools = {}
    for logn in range(12, 18):
        n = 1 << logn
        qs, seen = [], {}
        for i in range(40):
            bits = 30 + (i * 31) // 40
            k = seen.get(bits, 0)
            seen[bits] = k + 1
            qs.append(params.ntt_primes(n, bits, k + 1)[k])
        pools[logn] = [(ax.NTT(n, q, params.find_primitive_2n_root(n, q)), q) for q in qs]
    spec, total = [], 0
    for j in range(512):
        logn = int(rng.integers(12, 18))
        limbs = int(rng.integers(1, 41))
        inverse = bool(rng.integers(0, 2))
        lazy = bool(rng.integers(0, 2))
        spec.append((logn, limbs, inverse, lazy, int(rng.integers(0, 40))))
        total += limbs << logn
    src = torch.empty((total,), device="cuda", dtype=torch.int64)  # residues uniform in [0, q) per limb
    g = torch.Generator(device="cuda").manual_seed(12)
    out = torch.full_like(src, -1)
    ref = torch.full_like(src, -1)
    jobs, placed = [], []
    off = 0
    for logn, limbs, inverse, lazy, k in spec:
        n = 1 << logn
        of = (2 if inverse else 4) if lazy else 1
        for l in range(limbs):
            plan, q = pools[logn][(k + l) % 40]
            src[off:off + n] = torch.randint(0, q, (n,), generator=g, device="cuda", dtype=torch.int64)
            jobs.append((plan, out[off:off + n], src[off:off + n], inverse, of))
            placed.append((plan, q, off, n, inverse, of))
            off += n
    b = ax.Batch(jobs)
    b.launch()
    for plan, q, o, n, inverse, of in placed:
        (plan.ComputeInverse if inverse else plan.ComputeForward)(ref[o:o + n], src[o:o + n], 1, of)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    # a few limbs against the oracle: first, middle and last, one of each direction if present
    picks = {0, len(placed) // 2, len(placed) - 1}
    picks |= {next(i for i, x in enumerate(placed) if x[4]), next(i for i, x in enumerate(placed) if not x[4])}
    for i in sorted(picks):
        plan, q, o, n, inverse, of = placed[i]
        logn = n.bit_length() - 1
        psi = params.find_primitive_2n_root(n, q)
        t = O.Tables(n, q, psi)
        a = host(src[o:o + n])
        want = O.compute_inverse(a, t, 1, of) if inverse else O.compute_forward(a, t, 1, of)
        assert np.array_equal(host(out[o:o + n]), want), (i, logn, inverse, of)
    b.close()


def oracle_polymul(a, b, c, s, t, q):
    A = O.compute_forward(a, t, 1, 1)
    B = O.compute_forward(b, t, 1, 1)
    D = O.fma_mod(O.mul_mod(A, B, q, 1), s, c, q, 1)
    return O.compute_inverse(D, t, 1, 1)


def test_rns_product_full_size_fused_equals_call_sequence():
    torch = torch_mod()
    n, L, P = 1 << 15, 32, 128
    primes = params.rns_primes(n, L)
    psis = [params.find_primitive_2n_root(n, q) for q in primes]
    plans = [ax.NTT(n, q, p) for q, p in zip(primes, psis)]
    g = torch.Generator(device="cuda").manual_seed(4)
    a = torch.empty((P, L, n), dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    for l, q in enumerate(primes):
        for x in (a, b, c):
            x[:, l, :] = torch.randint(0, q, (P, n), generator=g, device="cuda", dtype=torch.int64)
    rng = np.random.default_rng(5)
    S = [[int(params.uniform(rng, 1, q)[0]) for q in primes] for _ in range(P)]
    outs = []
    for fused in (True, False):
        out = torch.empty_like(a)
        scr = torch.empty_like(a)
        pm = ax.Polymul([(plans[l], out[p, l], a[p, l], b[p, l], c[p, l], S[p][l], scr[p, l])
                         for p in range(P) for l in range(L)], fused=fused)
        if fused:
            assert pm.stacked_groups() == 1
        pm.launch()
        torch.cuda.synchronize()
        pm.close()
        del scr
        outs.append(out)
    assert torch.equal(outs[0], outs[1])
    for p, l in ((0, 0), (127, 31)):
        t = O.Tables(n, primes[l], psis[l])
        want = oracle_polymul(host(a[p, l]), host(b[p, l]), host(c[p, l]), S[p][l], t, primes[l])
        assert np.array_equal(host(outs[0][p, l]), want), (p, l)
