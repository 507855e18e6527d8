"""Batched NTT (many polynomials x RNS limbs, uneven degrees, one hipGraph)
bit-exact against the oracle, job by job."""
import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import dev, empty_dev, host, torch_mod

pytestmark = pytest.mark.gpu


def test_rns_batch_n65536_24_limbs_all_modes():
    """The 64 x 24 x 2^16 workload's shape at 4 polynomials: every limb of every
    polynomial, forward and inverse, lazy and normalized."""
    torch = torch_mod()
    n, L, P = 1 << 16, 24, 4
    primes = params.rns_primes(n, L)
    psis = [params.find_primitive_2n_root(n, q) for q in primes]
    plans = [ax.NTT(n, q, psi) for q, psi in zip(primes, psis)]
    tabs = [O.Tables(n, q, psi) for q, psi in zip(primes, psis)]
    rng = np.random.default_rng(1)
    host_in = np.stack([np.stack([params.uniform(rng, n, 4 * q) for q in primes]) for _ in range(P)])
    for inverse in (False, True):
        for lazy in (True, False):
            data = host_in if not inverse else np.stack(
                [np.stack([host_in[p, l] % (2 * primes[l]) for l in range(L)]) for p in range(P)])
            s = dev(data.reshape(-1)).view(P, L, n)
            out = torch.empty_like(s)
            of = (2 if inverse else 4) if lazy else 1
            jobs = [(plans[l], out[p, l], s[p, l], inverse, of) for p in range(P) for l in range(L)]
            b = ax.Batch(jobs)
            b.launch()
            got = host(out).reshape(P, L, n)
            for p in range(P):
                for l in range(L):
                    if inverse:
                        want = O.compute_inverse(data[p, l], tabs[l], 2, of)
                    else:
                        want = O.compute_forward(data[p, l], tabs[l], 4, of)
                    assert np.array_equal(got[p, l], want), (inverse, lazy, p, l)
            b.close()


def test_uneven_batch_graph_replays():
    """512-job style mix scaled down: N in 2^12..2^17, 1..6 limbs per job,
    30..60-bit moduli, both directions and factors; the captured graph is
    replayed three times with eager work in between, each replay checked."""
    torch = torch_mod()
    rng = np.random.default_rng(2024)
    jobs, refs, keep = [], [], []
    plan_cache = {}
    for j in range(40):
        n = 1 << int(rng.integers(12, 18))
        limbs = int(rng.integers(1, 7))
        inverse = bool(rng.integers(0, 2))
        lazy = bool(rng.integers(0, 2))
        of = (2 if inverse else 4) if lazy else 1
        for l in range(limbs):
            bits = int(rng.integers(30, 61))
            q = params.ntt_primes(n, bits, 1)[0]
            key = (n, q)
            if key not in plan_cache:
                psi = params.find_primitive_2n_root(n, q)
                plan_cache[key] = (ax.NTT(n, q, psi), O.Tables(n, q, psi))
            plan, tab = plan_cache[key]
            a = params.uniform(rng, n, (2 if inverse else 4) * q)
            s, o = dev(a), empty_dev(n)
            keep += [s, o]
            jobs.append((plan, o, s, inverse, of))
            want = O.compute_inverse(a, tab, 2, of) if inverse else O.compute_forward(a, tab, 4, of)
            refs.append((o, want))
    b = ax.Batch(jobs)
    for rep in range(3):
        for o, _ in refs:
            o.fill_(-1)
        b.launch()
        torch.cuda.synchronize()
        for i, (o, want) in enumerate(refs):
            assert np.array_equal(host(o), want), (rep, i)
        x = torch.arange(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert int(x) == (1 << 20) * ((1 << 20) - 1) // 2
    for mode in ("eager", "graph"):
        for o, _ in refs:
            o.fill_(-1)
        b.launch_eager() if mode == "eager" else b.launch_graph()
        for i, (o, want) in enumerate(refs):
            assert np.array_equal(host(o), want), (mode, i)
    b.close()


def test_single_group_batch_every_launch_mode():
    """A batch of one (degree, direction) group: launch issues its passes
    directly, launch_graph replays the captured graph, launch_eager issues
    them as plain launches; each gives the oracle's output on every replay,
    with unrelated eager work and synchronizations between replays."""
    torch = torch_mod()
    n = 1 << 14
    qs = params.rns_primes(n, 6)
    rng = np.random.default_rng(11)
    jobs, refs = [], []
    for q in qs:
        psi = params.find_primitive_2n_root(n, q)
        plan, tab = ax.NTT(n, q, psi), O.Tables(n, q, psi)
        for _ in range(3):
            a = params.uniform(rng, n, 4 * q)
            s, o = dev(a), empty_dev(n)
            jobs.append((plan, o, s, False, 1))
            refs.append((o, s, O.compute_forward(a, tab, 4, 1)))
    b = ax.Batch(jobs)
    assert b.kernel_count() >= 1
    for rep in range(3):
        for mode in ("launch", "graph", "eager"):
            for o, _, _ in refs:
                o.fill_(-1)
            {"launch": b.launch, "graph": b.launch_graph, "eager": b.launch_eager}[mode]()
            torch.cuda.synchronize()
            for i, (o, _, want) in enumerate(refs):
                assert np.array_equal(host(o), want), (rep, mode, i)
            x = torch.arange(1 << 18, device="cuda").sum()
            torch.cuda.synchronize()
            assert int(x) == (1 << 18) * ((1 << 18) - 1) // 2
    b.close()


def test_in_place_batch_jobs():
    n = 1 << 15
    qs = params.rns_primes(n, 8)
    rng = np.random.default_rng(3)
    jobs, refs = [], []
    keep = []
    for q in qs:
        psi = params.find_primitive_2n_root(n, q)
        plan, tab = ax.NTT(n, q, psi), O.Tables(n, q, psi)
        a = params.uniform(rng, n, q)
        d = dev(a)
        keep.append(plan)
        jobs.append((plan, d, d, False, 1))
        refs.append((d, O.compute_forward(a, tab, 1, 1)))
    b = ax.Batch(jobs)
    b.launch()
    for d, want in refs:
        assert np.array_equal(host(d), want)


@pytest.mark.parametrize("logn", [12, 13, 14, 15, 16, 17, 18, 19, 20])
def test_stacked_row_pass_every_config(logn):
    """Many jobs per plan: the ROW pass runs over stacks of same-plan jobs
    (one twiddle set per workgroup).  A plan's last partial stack is padded
    when at most one slot in eight stays empty (2 gpw - 1 jobs, gpw >= 8);
    otherwise its jobs (3 gpw + 1: one) run the per-limb ROW kernel beside
    the stacked one.  For every two-pass configuration, both directions and
    both output ranges, every job equals the plain per-limb call and the last
    job of each plan equals the oracle."""
    torch = torch_mod()
    n = 1 << logn
    qs = params.rns_primes(n, 3)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [ax.NTT(n, q, p) for q, p in zip(qs, psis)]
    tabs = [O.Tables(n, q, p) for q, p in zip(qs, psis)]
    rng = np.random.default_rng(logn)
    try:
        for ci in range(ax.config_count(n)):
            if ax.config_get(n, ci)["log_col"] == 0 or ax.config_get(n, ci)["row_groups"] < 2:
                continue  # no stacked ROW pass: single-pass, or one row per workgroup (nothing to share)
            ax.config_select(n, ci)
            gpw = ax.config_get(n, ci)["row_groups"]
            counts = [4 * gpw, 3 * gpw + 1, 2 * gpw - 1]  # whole stacks; one job left out; a padded stack
            owner = [pi for pi, c in enumerate(counts) for _ in range(c)]
            for inverse in (False, True):
                f_in = 2 if inverse else 4
                host_in = np.stack([params.uniform(rng, n, f_in * qs[pi]) for pi in owner])
                src = dev(host_in.reshape(-1)).view(len(owner), n)
                for lazy in (True, False):
                    of = f_in if lazy else 1
                    out = torch.full_like(src, -1)
                    b = ax.Batch([(plans[pi], out[i], src[i], inverse, of) for i, pi in enumerate(owner)])
                    assert b.stacked_groups() == 1, (logn, ci)
                    b.launch()
                    ref = torch.full_like(src, -1)
                    for i, pi in enumerate(owner):
                        call = plans[pi].ComputeInverse if inverse else plans[pi].ComputeForward
                        call(ref[i], src[i], f_in, of)
                    torch.cuda.synchronize()
                    assert torch.equal(out, ref), (logn, ci, inverse, lazy)
                    got = host(out).reshape(len(owner), n)
                    for pi in range(len(counts)):
                        i = owner.index(pi) + counts[pi] - 1  # last job of the plan
                        want = (O.compute_inverse(host_in[i], tabs[pi], f_in, of) if inverse
                                else O.compute_forward(host_in[i], tabs[pi], f_in, of))
                        assert np.array_equal(got[i], want), (logn, ci, inverse, lazy, pi)
                    b.close()
    finally:
        ax.config_select(n, -1)


@pytest.mark.parametrize("logn", [14, 16, 17])
def test_stacked_row_pass_on_8_byte_aligned_buffers(logn):
    """The stacked ROW pass (whole stacks, a padded stack and the per-limb
    rest) on buffers one word into an allocation, which take its 8-byte
    load and store paths: every two-pass configuration, both directions and
    both output ranges, every job equal to the plain call on a 16-byte
    aligned copy."""
    torch = torch_mod()
    n = 1 << logn
    qs = params.rns_primes(n, 3)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]
    rng = np.random.default_rng(100 + logn)
    try:
        for ci in range(ax.config_count(n)):
            if ax.config_get(n, ci)["log_col"] == 0 or ax.config_get(n, ci)["row_groups"] < 2:
                continue
            ax.config_select(n, ci)
            gpw = ax.config_get(n, ci)["row_groups"]
            counts = [2 * gpw, gpw + 1, 2 * gpw - 1]
            owner = [pi for pi, c in enumerate(counts) for _ in range(c)]
            jobs_n = len(owner)
            for inverse in (False, True):
                f_in = 2 if inverse else 4
                host_in = np.stack([params.uniform(rng, n, f_in * qs[pi]) for pi in owner])
                src = dev(host_in.reshape(-1)).view(jobs_n, n)
                base = torch.zeros(2 * jobs_n * n + 1, dtype=torch.int64, device="cuda")
                usrc = base[1:jobs_n * n + 1].view(jobs_n, n)
                uout = base[jobs_n * n + 1:].view(jobs_n, n)
                assert usrc.data_ptr() % 16 == 8 and uout.data_ptr() % 16 == 8
                for lazy in (True, False):
                    of = f_in if lazy else 1
                    usrc.copy_(src)
                    uout.fill_(-1)
                    b = ax.Batch([(plans[pi], uout[i], usrc[i], inverse, of) for i, pi in enumerate(owner)])
                    assert b.stacked_groups() == 1, (logn, ci)
                    b.launch()
                    ref = torch.full_like(src, -1)
                    for i, pi in enumerate(owner):
                        call = plans[pi].ComputeInverse if inverse else plans[pi].ComputeForward
                        call(ref[i], src[i], f_in, of)
                    torch.cuda.synchronize()
                    assert torch.equal(uout, ref), (logn, ci, inverse, lazy)
                    assert int(base[0].item()) == 0
                    b.close()
    finally:
        ax.config_select(n, -1)


def test_sparse_plans_do_not_stack():
    """One job per plan would leave most stack slots empty: the batch keeps
    the per-limb ROW pass."""
    torch = torch_mod()
    n = 1 << 15
    qs = params.rns_primes(n, 6)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]
    src = torch.zeros((6, n), dtype=torch.int64, device="cuda")
    out = torch.empty_like(src)
    b = ax.Batch([(plans[i], out[i], src[i], False, 1) for i in range(6)])
    assert b.stacked_groups() == 0
    b.launch()
    torch.cuda.synchronize()
    assert int(out.abs().sum()) == 0
    b.close()


def test_empty_batch_and_empty_product():
    torch = torch_mod()
    b = ax.Batch([])
    assert b.kernel_count() == 0 and b.stacked_groups() == 0
    b.launch()
    b.launch_graph()
    b.launch_eager()
    torch.cuda.synchronize()
    b.close()
    pm = ax.Polymul([])
    assert pm.stacked_groups() == 0
    pm.launch()
    torch.cuda.synchronize()
    pm.close()


def test_batch_rejects_host_pointers():
    """Batch jobs must be device memory (the plain calls stage host buffers;
    batches do not)."""
    torch = torch_mod()
    n = 1 << 12
    q = params.find_ntt_prime(n, 50)
    plan = ax.NTT(n, q, params.find_primitive_2n_root(n, q))
    d = torch.zeros(n, dtype=torch.int64, device="cuda")
    h = np.zeros(n, dtype=np.uint64)
    with pytest.raises(ax.AxhelError):
        ax.Batch([(plan, d, h.ctypes.data, False, 1)])
    with pytest.raises(ax.AxhelError):
        ax.Batch([(plan, h.ctypes.data, d, False, 1)])


def test_random_mixed_batches_equal_plain_calls():
    """Seeded random batches: degrees 2^1..2^17, moduli of 20..62 bits (both
    sides of the 2^60 schedule gate), both directions and output ranges,
    in-place and out-of-place jobs, buffers only 8-byte aligned (the batch
    then uses 8-byte accesses).  Every job equals the plain call; a sample
    equals the oracle."""
    torch = torch_mod()
    rng = np.random.default_rng(2026)
    for trial in range(3):
        specs, total = [], 0
        for _ in range(60):
            logn = int(rng.integers(1, 18))
            n = 1 << logn
            bits = int(rng.integers(max(20, logn + 3), 63))
            inverse = bool(rng.integers(0, 2))
            lazy = bool(rng.integers(0, 2))
            inplace = bool(rng.integers(0, 2))
            specs.append((n, bits, inverse, lazy, inplace))
            total += 2 * n + 1
        plans = {}
        for n, bits, *_ in specs:
            if (n, bits) not in plans:
                q = params.find_ntt_prime(n, bits)
                psi = params.find_primitive_2n_root(n, q)
                plans[(n, bits)] = (ax.NTT(n, q, psi), q, psi)
        base = torch.zeros(total + 1, dtype=torch.int64, device="cuda")
        arena = base[1:]  # one word in: every buffer is 8-byte, not 16-byte, aligned
        ref_arena = torch.zeros_like(arena)
        jobs, placed, off, inputs = [], [], 0, []
        for n, bits, inverse, lazy, inplace in specs:
            plan, q, psi = plans[(n, bits)]
            f_in = 2 if inverse else 4
            a = params.uniform(rng, n, f_in * q)
            src = arena[off:off + n]
            src.copy_(torch.from_numpy(a.view(np.int64)))
            dst = src if inplace else arena[off + n:off + 2 * n]
            rsrc = ref_arena[off:off + n]
            rsrc.copy_(src)
            rdst = rsrc if inplace else ref_arena[off + n:off + 2 * n]
            of = (f_in if lazy else 1)
            jobs.append((plan, dst, src, inverse, of))
            placed.append((plan, q, psi, n, inverse, of, dst, rdst, rsrc, a))
            off += 2 * n + 1
        b = ax.Batch(jobs)
        b.launch()
        for plan, q, psi, n, inverse, of, dst, rdst, rsrc, a in placed:
            (plan.ComputeInverse if inverse else plan.ComputeForward)(rdst, rsrc, 1, of)
        torch.cuda.synchronize()
        for i, (plan, q, psi, n, inverse, of, dst, rdst, rsrc, a) in enumerate(placed):
            assert torch.equal(dst, rdst), (trial, i, n, q.bit_length(), inverse, of)
        for i in range(0, len(placed), 7):
            plan, q, psi, n, inverse, of, dst, rdst, rsrc, a = placed[i]
            t = O.Tables(n, q, psi)
            want = O.compute_inverse(a, t, 1, of) if inverse else O.compute_forward(a, t, 1, of)
            assert np.array_equal(host(dst), want), (trial, i, n, q.bit_length(), inverse, of)
        b.close()
