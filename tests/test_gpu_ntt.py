"""Plain per-call NTT (ComputeForward / ComputeInverse drop-ins) on the GPU,
bit-exact against the oracle, lazy outputs included (tests/public/test_ntt.cpp
restated, extended to every compiled kernel configuration and N up to 2^20)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import NTT_DEGREES, MODULUS_BITS, SEED, dev, empty_dev, host, ntt_params, torch_mod

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def test_kat_n8_q17():
    k = GOLD["ntt_n8_q17"]
    ntt = ax.NTT(k["n"], k["q"], k["psi"])
    fwd, inv, ni = ntt.tables()
    assert fwd.reshape(-1, 2).tolist() == k["root_powers"]
    assert inv.reshape(-1, 2).tolist() == k["inv_root_powers"]
    assert ni.tolist() == k["inv_degree_modulo"]
    x = dev(k["input"])
    ntt.ComputeForward(x, x, 1, 1)
    assert host(x).tolist() == k["forward_normalized"]
    ntt.ComputeInverse(x, x, 1, 1)
    assert host(x).tolist() == k["input"]


def _check_all_modes(ntt, t, n, q, rng, label):
    a4 = O.MT19937_64(rng).random_vector(n, 4 * q)  # forward accepts [0, 4q)
    a2 = O.MT19937_64(rng + 1).random_vector(n, 2 * q)  # inverse accepts [0, 2q)
    d = dev(a4)
    r = empty_dev(n)
    ntt.ComputeForward(r, d, 4, 4)
    assert np.array_equal(host(r), O.compute_forward(a4, t, 4, 4)), ("fwd lazy", label)
    ntt.ComputeForward(r, d, 4, 1)
    assert np.array_equal(host(r), O.compute_forward(a4, t, 4, 1)), ("fwd norm", label)
    d = dev(a2)
    ntt.ComputeInverse(r, d, 2, 2)
    assert np.array_equal(host(r), O.compute_inverse(a2, t, 2, 2)), ("inv lazy", label)
    ntt.ComputeInverse(r, d, 2, 1)
    assert np.array_equal(host(r), O.compute_inverse(a2, t, 2, 1)), ("inv norm", label)


@pytest.mark.parametrize("logn", list(range(1, 21)))
def test_every_config_bit_exact(logn):
    n = 1 << logn
    try:
        for bits in (30, 50, 62):
            q, psi = ntt_params(n, bits)
            ntt = ax.NTT(n, q, psi)
            t = O.Tables(n, q, psi)
            for ci in range(ax.config_count(n)):
                ax.config_select(n, ci)
                _check_all_modes(ntt, t, n, q, SEED ^ (logn * 131 + bits + ci), (n, bits, ci))
    finally:
        ax.config_select(n, -1)


@pytest.mark.parametrize("n", NTT_DEGREES)
def test_round_trips_all_widths(n):
    """test_ntt.cpp:124-162 on the default configuration, every modulus width."""
    for bits in MODULUS_BITS:
        q, psi = ntt_params(n, bits)
        ntt = ax.NTT(n, q, psi)
        a = O.MT19937_64(SEED ^ n ^ bits).random_vector(n, q)
        x = dev(a)
        ntt.ComputeForward(x, x, 1, 1)
        ntt.ComputeInverse(x, x, 1, 1)
        assert np.array_equal(host(x), a), (n, bits)
        x = dev(a)
        ntt.ComputeForward(x, x, 1, 4)
        ax.EltwiseReduceMod(x, x, n, q, 4, 2)
        ntt.ComputeInverse(x, x, 2, 1)
        assert np.array_equal(host(x), a), ("lazy", n, bits)


def test_out_of_place_matches_in_place_and_host_pointers():
    for n in (8, 128, 1024, 8192, 32768, 65536, 131072, 1 << 20):
        q, psi = ntt_params(n, 62)
        ntt = ax.NTT(n, q, psi)
        a = O.MT19937_64(SEED ^ n).random_vector(n, q)
        ip = dev(a)
        ntt.ComputeForward(ip, ip, 1, 1)
        src = dev(a)
        out = empty_dev(n)
        ntt.ComputeForward(out, src, 1, 1)
        assert np.array_equal(host(out), host(ip))
        assert np.array_equal(host(src), a)  # operand untouched out of place
        h = a.copy()
        ntt.ComputeForward(h, h, 1, 1)  # host buffer, staged
        assert np.array_equal(h, host(ip))
        ho = np.zeros(n, dtype=np.uint64)
        ntt.ComputeInverse(ho, h, 1, 1)
        assert np.array_equal(ho, a)


def test_external_tables_constructor():
    k = GOLD["ntt_n8_q17"]
    ext = ax.NTT(8, 17, root_powers=np.array(k["root_powers"], dtype=np.uint64).ravel(),
                 inv_root_powers=np.array(k["inv_root_powers"], dtype=np.uint64).ravel(),
                 inv_degree_modulo=k["inv_degree_modulo"])
    x = dev(k["input"])
    ext.ComputeForward(x, x, 1, 1)
    assert host(x).tolist() == k["forward_normalized"]
    ext.ComputeInverse(x, x, 1, 1)
    assert host(x).tolist() == k["input"]
    n = 4096
    q, psi = ntt_params(n, 50)
    t = O.Tables(n, q, psi)
    ext = ax.NTT(n, q, root_powers=t.fwd, inv_root_powers=t.inv, inv_degree_modulo=t.ninv)
    assert ext.Degree() == n and ext.Modulus() == q
    a = O.MT19937_64(3).random_vector(n, 4 * q)
    x = dev(a)
    ext.ComputeForward(x, x, 1, 4)
    assert np.array_equal(host(x), O.compute_forward(a, t, 1, 4))


def test_external_tables_are_borrowed():
    """ntt.hpp:34-37: the plan borrows the caller's tables, so a call sees
    them as they are when it is made.  The caller rewrites its arrays in
    place with the tables of another primitive 2N-th root (psi^3), and back;
    each call equals the oracle on the tables it saw, forward and inverse
    (the inverse's scaled final root follows the last inverse entry)."""
    n = 2048
    q, psi = ntt_params(n, 55)
    t1 = O.Tables(n, q, psi)
    t3 = O.Tables(n, q, pow(psi, 3, q))
    fwd = t1.fwd.copy()
    inv = t1.inv.copy()
    ext = ax.NTT(n, q, root_powers=fwd, inv_root_powers=inv, inv_degree_modulo=t1.ninv)
    a4 = O.MT19937_64(5).random_vector(n, 4 * q)
    a2 = O.MT19937_64(6).random_vector(n, 2 * q)
    r = empty_dev(n)
    for t in (t1, t3, t1, t3):
        fwd[:] = t.fwd
        inv[:] = t.inv
        for of in (4, 1):
            ext.ComputeForward(r, dev(a4), 4, of)
            assert np.array_equal(host(r), O.compute_forward(a4, t, 4, of)), (t is t3, of)
        for of in (2, 1):
            ext.ComputeInverse(r, dev(a2), 2, of)
            assert np.array_equal(host(r), O.compute_inverse(a2, t, 2, of)), (t is t3, of)
    # a change to one table only: the forward table back to psi, the inverse kept at psi^3
    fwd[:] = t1.fwd
    ext.ComputeForward(r, dev(a4), 4, 1)
    assert np.array_equal(host(r), O.compute_forward(a4, t1, 4, 1))
    ext.ComputeInverse(r, dev(a2), 2, 1)
    assert np.array_equal(host(r), O.compute_inverse(a2, t3, 2, 1))


def test_special_patterns():
    """test_ntt.cpp:15-30 patterns, lazy and normalized, against the oracle."""
    for n in (8, 4096, 65536):
        q, psi = ntt_params(n, 60)
        ntt = ax.NTT(n, q, psi)
        t = O.Tables(n, q, psi)
        pats = [np.zeros(n, np.uint64), np.eye(1, n, 0, dtype=np.uint64)[0], np.ones(n, np.uint64),
                np.full(n, q - 1, np.uint64), np.array([(q - 1) * (i & 1) for i in range(n)], np.uint64),
                np.array([q - 1 if i == n - 1 else 0 for i in range(n)], np.uint64),
                np.full(n, 4 * q - 1, np.uint64)]
        for a in pats:
            r = empty_dev(n)
            ntt.ComputeForward(r, dev(a), 1, 4)
            assert np.array_equal(host(r), O.compute_forward(a, t, 1, 4))
            ntt.ComputeForward(r, dev(a), 1, 1)
            assert np.array_equal(host(r), O.compute_forward(a, t, 1, 1))
            b = np.minimum(a, 2 * q - 1)
            ntt.ComputeInverse(r, dev(b), 2, 2)
            assert np.array_equal(host(r), O.compute_inverse(b, t, 2, 2))


def test_invalid_plans_are_rejected():
    with pytest.raises(ax.AxhelError):
        ax.NTT(12, 97, 3)  # not a power of two
    with pytest.raises(ax.AxhelError):
        ax.NTT(1 << 21, 0x3FFFFFFFFFF00001, 3)  # above the supported degree


def test_library_loaded_before_torch_leaves_torch_a_gpu():
    """Creating a plan (HIP calls) before torch is imported must not leave the
    process with a second HIP runtime that hides the GPU from torch."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; import axhel_amd as ax; from axhel_amd import params; n = 1024; "
            "q = params.find_ntt_prime(n, 50); p = ax.NTT(n, q, params.find_primitive_2n_root(n, q)); "
            "assert 'torch' in sys.modules; import torch; assert torch.cuda.is_available(); print('ok')")
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-2000:]


@pytest.mark.parametrize("logn", [1, 2, 3, 5, 8, 12, 13, 14, 16, 17])
def test_device_generated_tables_equal_the_oracle(logn):
    """Plan tables are generated on the device (powers by square-and-multiply,
    Shoup quotients by long division): every word equals the oracle's tables
    (ntt-tables.cpp:81-140 restated), for moduli of 20..62 bits, and a root
    given above q gives the same tables as its residue."""
    n = 1 << logn
    for bits in (20, 41, 50, 62):
        if (1 << bits) <= 2 * n:
            continue
        q = params.find_ntt_prime(n, bits)
        psi = params.find_primitive_2n_root(n, q)
        t = O.Tables(n, q, psi)
        for root in (psi, psi + q):
            fwd, inv, ni = ax.NTT(n, q, root).tables()
            assert np.array_equal(fwd, t.fwd), (logn, bits)
            assert np.array_equal(inv, t.inv), (logn, bits)
            assert np.array_equal(ni, t.ninv), (logn, bits)


@pytest.mark.parametrize("logn", [1, 4, 12, 13, 14, 16, 17])
def test_relaxed_schedule_boundary_moduli(logn):
    """A normalized forward with every modulus below 2^60 runs on the relaxed
    butterfly schedule (intermediate words up to 16q); at or above 2^60 it
    runs the reference's schedule.  The largest NTT prime below 2^60 with
    inputs at the top of the domain (every word 4q - 1) and random inputs,
    and the largest below 2^61, equal the oracle through plain calls, a batch
    group that mixes both (exact schedule for the group), a batch group of
    the 60-bit prime alone (relaxed) and an RNS context over both."""
    torch = pytest.importorskip("torch")
    n = 1 << logn
    qs = [params.find_ntt_prime(n, 60), params.find_ntt_prime(n, 61)]
    assert qs[0] < 1 << 60 <= qs[1]
    psis = [params.find_primitive_2n_root(n, q) for q in qs]
    plans = [ax.NTT(n, q, p) for q, p in zip(qs, psis)]
    tabs = [O.Tables(n, q, p) for q, p in zip(qs, psis)]
    ins = []
    for l, q in enumerate(qs):
        ins.append([np.full(n, 4 * q - 1, np.uint64), O.MT19937_64(SEED + l).random_vector(n, 4 * q)])
    want = [[O.compute_forward(a, tabs[l], 1, 1) for a in ins[l]] for l in range(2)]
    for l in range(2):
        for k, a in enumerate(ins[l]):
            r = empty_dev(n)
            plans[l].ComputeForward(r, dev(a), 1, 1)
            assert np.array_equal(host(r), want[l][k]), ("plain", l, k)
    src = torch.from_numpy(np.stack([np.stack(ins[0]), np.stack(ins[1])]).view(np.int64)).cuda()  # [limb][k][n]
    out = torch.full_like(src, -1)
    mixed = ax.Batch([(plans[l], out[l, k], src[l, k], False, 1) for l in range(2) for k in range(2)])
    mixed.launch()
    torch.cuda.synchronize()
    for l in range(2):
        for k in range(2):
            assert np.array_equal(host(out[l, k]), want[l][k]), ("mixed batch", l, k)
    out.fill_(-1)
    alone = ax.Batch([(plans[0], out[0, k], src[0, k], False, 1) for k in range(2)])
    alone.launch()
    torch.cuda.synchronize()
    for k in range(2):
        assert np.array_equal(host(out[0, k]), want[0][k]), ("60-bit batch", k)
    mixed.close()
    alone.close()
    ctx = ax.RNS(n, qs, psis)
    poly = src.permute(1, 0, 2).contiguous()  # [k][limb][n]
    res = torch.full_like(poly, -1)
    ctx.ntt(res, poly, 2, False, 1)
    torch.cuda.synchronize()
    for l in range(2):
        for k in range(2):
            assert np.array_equal(host(res[k, l]), want[l][k]), ("rns", l, k)
    ctx.close()


def test_host_pointer_calls_reuse_pooled_staging_buffers():
    """Host-pointer calls stage through pooled device buffers, handed back
    after each call's synchronisation: calls of changing sizes in a row, in
    place and out of place, equal the oracle; then two threads issue
    host-pointer calls at once (distinct buffers from the shared pool)."""
    import threading
    sizes = [1 << 10, 1 << 16, 1 << 12, 1 << 17, 1 << 16, 1 << 10, 1 << 17]
    plans = {}
    for i, n in enumerate(sizes):
        if n not in plans:
            q, psi = ntt_params(n, 50)
            plans[n] = (ax.NTT(n, q, psi), O.Tables(n, q, psi), q)
        ntt, t, q = plans[n]
        a = O.MT19937_64(SEED + i).random_vector(n, 4 * q)
        want = O.compute_forward(a, t, 1, 1)
        out = np.zeros(n, dtype=np.uint64)
        ntt.ComputeForward(out, a, 1, 1)
        assert np.array_equal(out, want), (i, n)
        h = a.copy()
        ntt.ComputeForward(h, h, 1, 4)
        assert np.array_equal(h, O.compute_forward(a, t, 1, 4)), (i, n)
    n = 1 << 16
    ntt, t, q = plans[n]
    inputs = [O.MT19937_64(SEED + 100 + k).random_vector(n, 4 * q) for k in range(2)]
    wants = [O.compute_forward(a, t, 1, 1) for a in inputs]
    errors = []

    def worker(k):
        try:
            for _ in range(20):
                out = np.zeros(n, dtype=np.uint64)
                ntt.ComputeForward(out, inputs[k], 1, 1)
                if not np.array_equal(out, wants[k]):
                    errors.append(k)
                    return
        except Exception as e:  # reported below
            errors.append(repr(e))

    ths = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors


@pytest.mark.parametrize("logn", [1, 4, 10, 13, 16, 17])
def test_plain_calls_on_8_byte_aligned_buffers(logn):
    """Only 8-byte alignment is guaranteed (the reference's UnalignedOffsets
    cases, test_eltwise.cpp:518-561): operands and results one word into an
    allocation take the 8-byte access path of every compiled configuration,
    in place and out of place, and equal the oracle."""
    torch = torch_mod()
    n = 1 << logn
    q, psi = ntt_params(n, 55)
    ntt = ax.NTT(n, q, psi)
    t = O.Tables(n, q, psi)
    a4 = O.MT19937_64(SEED + logn).random_vector(n, 4 * q)
    a2 = O.MT19937_64(SEED + logn + 1).random_vector(n, 2 * q)
    base = torch.zeros(2 * n + 2, dtype=torch.int64, device="cuda")
    src, dst = base[1:n + 1], base[n + 1:2 * n + 1]
    assert src.data_ptr() % 16 == 8 and dst.data_ptr() % 16 == 8
    try:
        for ci in range(ax.config_count(n)):
            ax.config_select(n, ci)
            for inverse, a, f in ((False, a4, 4), (True, a2, 2)):
                call = ntt.ComputeInverse if inverse else ntt.ComputeForward
                want_f = O.compute_inverse if inverse else O.compute_forward
                for of in (f, 1):
                    src.copy_(torch.from_numpy(a.view(np.int64)))
                    call(dst, src, f, of)
                    assert np.array_equal(host(dst), want_f(a, t, f, of)), (ci, inverse, of, "out of place")
                    call(src, src, f, of)
                    assert np.array_equal(host(src), want_f(a, t, f, of)), (ci, inverse, of, "in place")
        assert bool((base[0] == 0).item()) and bool((base[-1] == 0).item())  # neighbours untouched
    finally:
        ax.config_select(n, -1)


def test_rns_calls_on_8_byte_aligned_buffers():
    """An RNS context's whole-polynomial calls on arrays one word into an
    allocation equal the same calls on 16-byte-aligned arrays."""
    torch = torch_mod()
    n, L, count = 1 << 14, 3, 4
    qs = params.rns_primes(n, L)
    ctx = ax.RNS(n, qs)
    qv = torch.tensor(qs, device="cuda").view(1, L, 1)
    words = count * L * n
    src = torch.remainder(torch.randint(0, 1 << 62, (count, L, n), dtype=torch.int64, device="cuda"), qv)
    base = torch.zeros(2 * words + 1, dtype=torch.int64, device="cuda")
    usrc, udst = base[1:words + 1].view(count, L, n), base[words + 1:].view(count, L, n)
    usrc.copy_(src)
    assert usrc.data_ptr() % 16 == 8
    for inverse in (False, True):
        want = torch.empty_like(src)
        ctx.ntt(want, src, count, inverse, 1)
        ctx.ntt(udst, usrc, count, inverse, 1)
        torch.cuda.synchronize()
        assert torch.equal(udst, want), inverse
    want = torch.empty_like(src)
    ctx.eltwise(ax.OP_MUL, want, src, src, None, count, 1, 1)
    ctx.eltwise(ax.OP_MUL, udst, usrc, usrc, None, count, 1, 1)
    torch.cuda.synchronize()
    assert torch.equal(udst, want) and bool((base[0] == 0).item())
    ctx.close()
