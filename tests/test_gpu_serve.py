"""Serving loop on the GPU: every request's result is bit-identical run alone
and run in a concurrent mix (continuous batching), and equal to the oracle."""
import threading

import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import dev, empty_dev, host, torch_mod

pytestmark = pytest.mark.gpu


def make_requests(server, rng, count):
    """Mixed NTT / INTT / element-wise requests, N in 2^13..2^17, 1..6 limbs."""
    reqs = []
    cache = {}
    for i in range(count):
        n = 1 << int(rng.integers(13, 18))
        limbs = int(rng.integers(1, 7))
        kind = ["fwd", "inv", "mul", "fma", "add", "red"][i % 6]
        qs = params.rns_primes(n, limbs, bit_schedule=[50, 55, 60])
        psis = [params.find_primitive_2n_root(n, q) for q in qs]
        for q, psi in zip(qs, psis):
            if (n, q) not in cache:
                cache[(n, q)] = O.Tables(n, q, psi)
        bound = {"fwd": 4, "inv": 2, "red": 4}.get(kind, 1)
        a = np.concatenate([params.uniform(rng, n, bound * q) for q in qs])
        b = np.concatenate([params.uniform(rng, n, q) for q in qs])
        sc = [int(params.uniform(rng, 1, q)[0]) for q in qs]
        want = []
        for l, q in enumerate(qs):
            t = cache[(n, q)]
            al, bl = a[l * n:(l + 1) * n], b[l * n:(l + 1) * n]
            if kind == "fwd":
                want.append(O.compute_forward(al, t, 4, 4 if l % 2 else 1))
            elif kind == "inv":
                want.append(O.compute_inverse(al, t, 2, 1))
            elif kind == "mul":
                want.append(O.mul_mod(al, bl, q, 1))
            elif kind == "fma":
                want.append(O.fma_mod(al, sc[l], bl, q, 1))
            elif kind == "add":
                want.append(O.add_mod(al, bl, q))
            else:
                want.append(O.reduce_mod(al, q, 4, 1))
        plans = [server.plan(n, q, psi) for q, psi in zip(qs, psis)] if kind in ("fwd", "inv") else None
        reqs.append(dict(kind=kind, n=n, limbs=limbs, qs=qs, plans=plans, a=a, b=b, sc=sc,
                         want=np.concatenate(want)))
    return reqs


def submit(server, R, result, a, b):
    kind = R["kind"]
    if kind == "fwd":
        # lazy output on odd limbs is not expressible per limb: forward requests
        # use one factor; the mixed-factor reference is split below
        return None
    common = dict(degree=R["n"], limbs=R["limbs"], result=result, operand1=a)
    if kind == "inv":
        return server.submit(ax.OP_NTT_INVERSE, operand2=None, plans=R["plans"], output_mod_factor=1, **common)
    if kind == "mul":
        return server.submit(ax.OP_MUL, operand2=b, moduli=R["qs"], input_mod_factor=1, **common)
    if kind == "fma":
        return server.submit(ax.OP_FMA, operand2=b, moduli=R["qs"], scalars=R["sc"], input_mod_factor=1, **common)
    if kind == "add":
        return server.submit(ax.OP_ADD, operand2=b, moduli=R["qs"], **common)
    return server.submit(ax.OP_REDUCE, operand2=None, moduli=R["qs"], input_mod_factor=4, output_mod_factor=1,
                         **common)


def submit_fwd(server, R, result, a, lazy):
    return server.submit(ax.OP_NTT_FORWARD, degree=R["n"], limbs=R["limbs"], result=result, operand1=a,
                         plans=R["plans"], output_mod_factor=4 if lazy else 1)


def oracle_fwd(R, lazy):
    n = R["n"]
    out = []
    for l, q in enumerate(R["qs"]):
        t = O.Tables(n, q, params.find_primitive_2n_root(n, q))
        out.append(O.compute_forward(R["a"][l * n:(l + 1) * n], t, 4, 4 if lazy else 1))
    return np.concatenate(out)


def test_alone_equals_mix_equals_oracle():
    torch = torch_mod()
    rng = np.random.default_rng(77)
    server = ax.Server(slots=3)
    reqs = make_requests(server, rng, 30)
    for R in reqs:
        if R["kind"] == "fwd":
            R["lazy"] = bool(rng.integers(0, 2))
            R["want"] = oracle_fwd(R, R["lazy"])
        R["da"], R["db"] = dev(R["a"]), dev(R["b"])

    def run(R):
        out = empty_dev(R["n"] * R["limbs"])
        if R["kind"] == "fwd":
            st, t = submit_fwd(server, R, out, R["da"], R["lazy"])
        else:
            st, t = submit(server, R, out, R["da"], R["db"])
        assert st == ax.OK, (R["kind"], st)
        return out, t

    # alone: one request at a time
    alone = []
    for R in reqs:
        out, t = run(R)
        assert server.wait(t) == ax.OK
        alone.append(host(out))
    # mix: every request in flight at once, submitted from four threads
    outs = [None] * len(reqs)
    tickets = [None] * len(reqs)

    def client(idx):
        for i in idx:
            outs[i], tickets[i] = run(reqs[i])

    ths = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    for t in tickets:
        assert server.wait(t) == ax.OK
    stt = server.stats()
    assert stt["failed"] == 0 and stt["completed"] == 2 * len(reqs)
    for i, R in enumerate(reqs):
        mix = host(outs[i])
        assert np.array_equal(alone[i], mix), (i, R["kind"])
        assert np.array_equal(mix, R["want"]), (i, R["kind"], R["n"])
    server.close()


def test_host_buffers_are_staged_through_the_pool():
    rng = np.random.default_rng(5)
    server = ax.Server(slots=2, memory_budget_bytes=64 << 20)
    n, limbs = 1 << 14, 3
    qs = params.rns_primes(n, limbs)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [server.plan(n, q, p) for q, p in zip(qs, psis)]
    a = np.concatenate([params.uniform(rng, n, q) for q in qs])
    r = np.zeros_like(a)
    st, t = server.submit(ax.OP_NTT_FORWARD, n, limbs, r, a, plans=plans, output_mod_factor=1)
    assert st == ax.OK and server.wait(t) == ax.OK
    for l, (q, p) in enumerate(zip(qs, psis)):
        assert np.array_equal(r[l * n:(l + 1) * n], O.compute_forward(a[l * n:(l + 1) * n], O.Tables(n, q, p)))
    x = r.copy()
    st, t = server.submit(ax.OP_NTT_INVERSE, n, limbs, x, x, plans=plans)  # in place on a host buffer
    assert st == ax.OK and server.wait(t) == ax.OK
    assert np.array_equal(x, a)
    assert server.stats()["bytes_in_use"] == 0
    server.close()


def test_requests_on_8_byte_aligned_buffers():
    """Operands and results one word into an allocation (only 8-byte
    alignment is guaranteed) take the servers' 8-byte access path; every kind
    of request still equals the oracle, in place and out of place."""
    torch = torch_mod()
    rng = np.random.default_rng(91)
    server = ax.Server(slots=2)
    reqs = make_requests(server, rng, 12)
    for i, R in enumerate(reqs):
        if R["kind"] == "fwd":
            R["lazy"] = bool(i % 4 == 0)
            R["want"] = oracle_fwd(R, R["lazy"])
        words = R["n"] * R["limbs"]
        base = torch.zeros(3 * words + 1, dtype=torch.int64, device="cuda")
        da, db, out = base[1:words + 1], base[words + 1:2 * words + 1], base[2 * words + 1:]
        da.copy_(dev(R["a"]))
        db.copy_(dev(R["b"]))
        assert da.data_ptr() % 16 == 8 and out.data_ptr() % 16 == 8
        inplace = i % 3 == 2
        dst = da if inplace else out
        if R["kind"] == "fwd":
            st, t = submit_fwd(server, R, dst, da, R["lazy"])
        else:
            st, t = submit(server, R, dst, da, db)
        assert st == ax.OK and server.wait(t) == ax.OK, (i, R["kind"])
        assert np.array_equal(host(dst), R["want"]), (i, R["kind"], R["n"], inplace)
        assert bool((base[0] == 0).item())
    server.close()


def test_requests_after_a_failed_batch_are_correct():
    """Every second device batch fails after its host operands' staging
    copies were queued (injected).  The server drains that stream before its
    staging blocks go back to the pool, so the next request, which reuses
    them, still equals the oracle; the failed requests end NO_MEMORY."""
    rng = np.random.default_rng(17)
    server = ax.Server(slots=1, memory_budget_bytes=256 << 20, fail_every_batch=2)
    n, limbs = 1 << 15, 4
    qs = params.rns_primes(n, limbs)
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [server.plan(n, q, p) for q, p in zip(qs, psis)]
    tabs = [O.Tables(n, q, p) for q, p in zip(qs, psis)]
    got = []
    for i in range(6):  # one request per batch: submit, then wait
        a = np.concatenate([params.uniform(rng, n, q) for q in qs])
        r = np.zeros_like(a)
        st, t = server.submit(ax.OP_NTT_FORWARD, n, limbs, r, a, plans=plans, output_mod_factor=1)
        assert st == ax.OK
        res = server.wait(t)
        got.append(res)
        if res == ax.OK:
            for l in range(limbs):
                assert np.array_equal(r[l * n:(l + 1) * n], O.compute_forward(a[l * n:(l + 1) * n], tabs[l])), (i, l)
    assert got == [ax.OK, ax.NO_MEMORY] * 3, got
    stt = server.stats()
    assert stt["failed"] == 3 and stt["bytes_in_use"] == 0 and stt["descriptor_bytes"] > 0
    server.close()


def test_wide_requests_alone_equal_mix_equal_oracle():
    """The serving mix at its configured width: 1..48 limbs per request and
    N in 2^13..2^17, so single requests (up to 6.3 M words) and whole batches
    cross the 2^21-word switch to the throughput configuration.  Each
    result is the same alone and in the concurrent mix, and its first and
    last limb equal the oracle."""
    torch = torch_mod()
    rng = np.random.default_rng(2026)
    server = ax.Server(slots=4)
    kinds = ["fwd", "inv", "mul", "fwd", "inv", "fma"]
    reqs, cache = [], {}
    for i in range(18):
        n = 1 << int(rng.integers(13, 18))
        limbs = 48 if i < 2 else int(rng.integers(1, 49))
        kind = kinds[i % len(kinds)]
        qs = params.rns_primes(n, limbs)
        psis = [params.find_primitive_2n_root(n, q) for q in qs]
        bound = {"fwd": 4, "inv": 2}.get(kind, 1)
        a = np.concatenate([params.uniform(rng, n, bound * q) for q in qs])
        b = np.concatenate([params.uniform(rng, n, q) for q in qs])
        sc = [int(params.uniform(rng, 1, q)[0]) for q in qs]
        lazy = bool(i % 4 == 3)
        want = {}
        for l in sorted({0, limbs - 1}):
            q, psi = qs[l], psis[l]
            if kind in ("fwd", "inv") and (n, q) not in cache:
                cache[(n, q)] = O.Tables(n, q, psi)
            al, bl = a[l * n:(l + 1) * n], b[l * n:(l + 1) * n]
            if kind == "fwd":
                want[l] = O.compute_forward(al, cache[(n, q)], 4, 4 if lazy else 1)
            elif kind == "inv":
                want[l] = O.compute_inverse(al, cache[(n, q)], 2, 2 if lazy else 1)
            elif kind == "mul":
                want[l] = O.mul_mod(al, bl, q, 1)
            else:
                want[l] = O.fma_mod(al, sc[l], bl, q, 1)
        plans = [server.plan(n, q, psi) for q, psi in zip(qs, psis)] if kind in ("fwd", "inv") else None
        reqs.append(dict(kind=kind, n=n, limbs=limbs, qs=qs, plans=plans, sc=sc, lazy=lazy, want=want,
                         da=dev(a), db=dev(b)))
    assert max(R["n"] * R["limbs"] for R in reqs) > 1 << 21

    def run(R):
        out = empty_dev(R["n"] * R["limbs"])
        common = dict(degree=R["n"], limbs=R["limbs"], result=out, operand1=R["da"])
        if R["kind"] == "fwd":
            st, t = server.submit(ax.OP_NTT_FORWARD, plans=R["plans"], output_mod_factor=4 if R["lazy"] else 1,
                                  **common)
        elif R["kind"] == "inv":
            st, t = server.submit(ax.OP_NTT_INVERSE, plans=R["plans"], output_mod_factor=2 if R["lazy"] else 1,
                                  **common)
        elif R["kind"] == "mul":
            st, t = server.submit(ax.OP_MUL, operand2=R["db"], moduli=R["qs"], input_mod_factor=1, **common)
        else:
            st, t = server.submit(ax.OP_FMA, operand2=R["db"], moduli=R["qs"], scalars=R["sc"], input_mod_factor=1,
                                  **common)
        assert st == ax.OK, (R["kind"], st)
        return out, t

    alone = []
    for R in reqs:
        out, t = run(R)
        assert server.wait(t) == ax.OK
        alone.append(host(out))
    outs, tickets = [None] * len(reqs), [None] * len(reqs)

    def client(idx):
        for i in idx:
            outs[i], tickets[i] = run(reqs[i])

    ths = [threading.Thread(target=client, args=(list(range(k, len(reqs), 3)),)) for k in range(3)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    for t in tickets:
        assert server.wait(t) == ax.OK
    for i, R in enumerate(reqs):
        mix = host(outs[i])
        assert np.array_equal(alone[i], mix), (i, R["kind"], R["n"], R["limbs"])
        n = R["n"]
        for l, w in R["want"].items():
            assert np.array_equal(mix[l * n:(l + 1) * n], w), (i, R["kind"], n, R["limbs"], l)
    assert server.stats()["failed"] == 0
    server.close()
