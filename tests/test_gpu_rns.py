"""RNS context: calls over count polynomials x L limbs stored [count][L][N]
equal the reference's per-limb calls (the caller's limb loop,
integration/seal/SEAL-4.4.5-AXHEL.patch:185-226), limb by limb, bit for bit;
sampled limbs equal the oracle."""
import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from oracle import oracle as O
from tests.helpers import host, torch_mod

pytestmark = pytest.mark.gpu


def setup(n, L, count, seed, bound_factor=1):
    torch = torch_mod()
    qs = params.rns_primes(n, L)
    ctx = ax.RNS(n, qs)
    rng = np.random.default_rng(seed)
    data = np.stack([np.stack([params.uniform(rng, n, bound_factor * q) for q in qs]) for _ in range(count)])
    src = torch.from_numpy(data.view(np.int64).copy()).cuda()
    return torch, qs, ctx, data, src


@pytest.mark.parametrize("logn,count", [(12, 3), (12, 40), (16, 5), (16, 24)])
def test_rns_ntt_equals_per_limb_calls(logn, count):
    n, L = 1 << logn, 6
    for inverse in (False, True):
        f_in = 2 if inverse else 4
        torch, qs, ctx, data, src = setup(n, L, count, logn + count + inverse, f_in)
        plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]
        for lazy in (True, False):
            of = f_in if lazy else 1
            out = torch.full_like(src, -1)
            ctx.ntt(out, src, count, inverse, of)
            ref = torch.full_like(src, -1)
            for p in range(count):
                for l in range(L):
                    (plans[l].ComputeInverse if inverse else plans[l].ComputeForward)(ref[p, l], src[p, l], 1, of)
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (logn, count, inverse, lazy)
            t = O.Tables(n, qs[L - 1], params.find_primitive_2n_root(n, qs[L - 1]))
            a = data[count - 1, L - 1]
            want = O.compute_inverse(a, t, f_in, of) if inverse else O.compute_forward(a, t, f_in, of)
            assert np.array_equal(host(out[count - 1, L - 1]), want)
        # in place
        buf = src.clone()
        ctx.ntt(buf, buf, count, inverse, 1)
        ctx.ntt(out, src, count, inverse, 1)
        torch.cuda.synchronize()
        assert torch.equal(buf, out)
        ctx.close()


def test_rns_default_roots_follow_the_reference_rule():
    n = 1 << 10
    qs = params.rns_primes(n, 3)
    ctx = ax.RNS(n, qs)
    for l, q in enumerate(qs):
        plan = ax.NTT(n, q, params.find_primitive_2n_root(n, q))
        fwd, inv, ni = plan.tables()
        got = np.zeros(2 * n, dtype=np.uint64)
        assert ax.lib().axhel_hip_ntt_plan_tables(ctx.plan_handle(l), got.ctypes.data, None, None) == 0
        assert np.array_equal(got, fwd), l
    ctx.close()


@pytest.mark.parametrize("logn", [12, 15])
def test_rns_eltwise_equals_per_limb_calls(logn):
    n, L, count = 1 << logn, 5, 7
    torch, qs, ctx, a_np, a = setup(n, L, count, 11, 4)
    rng = np.random.default_rng(12)
    b_np = np.stack([np.stack([params.uniform(rng, n, 4 * q) for q in qs]) for _ in range(count)])
    b = torch.from_numpy(b_np.view(np.int64).copy()).cuda()
    lo = torch.remainder(a, torch.tensor(qs, device="cuda").view(1, L, 1))  # canonical operands
    lo_b = torch.remainder(b, torch.tensor(qs, device="cuda").view(1, L, 1))
    scalars = [int(params.uniform(rng, 1, 8 * q)[0]) for q in qs]
    small = [int(params.uniform(rng, 1, q)[0]) for q in qs]  # vector-scalar add/sub take a scalar below q
    cases = [
        (ax.OP_ADD, lo, lo_b, None, 1, 1, lambda r, x, y, q, l: ax.EltwiseAddMod(r, x, y, n, q)),
        (ax.OP_SUB, lo, lo_b, None, 1, 1, lambda r, x, y, q, l: ax.EltwiseSubMod(r, x, y, n, q)),
        (ax.OP_ADD, lo, None, small, 1, 1, lambda r, x, y, q, l: ax.EltwiseAddMod(r, x, small[l], n, q)),
        (ax.OP_SUB, lo, None, small, 1, 1, lambda r, x, y, q, l: ax.EltwiseSubMod(r, x, small[l], n, q)),
        (ax.OP_MUL, lo, lo_b, None, 1, 1, lambda r, x, y, q, l: ax.EltwiseMulMod(r, x, y, n, q, 1)),
        (ax.OP_MUL, a, b, None, 4, 1, lambda r, x, y, q, l: ax.EltwiseMulMod(r, x, y, n, q, 4)),
        (ax.OP_FMA, a, b, scalars, 8, 1, lambda r, x, y, q, l: ax.EltwiseFMAMod(r, x, scalars[l], y, n, q, 8)),
        (ax.OP_FMA, lo, None, scalars, 8, 1, lambda r, x, y, q, l: ax.EltwiseFMAMod(r, x, scalars[l], None, n, q, 8)),
        (ax.OP_REDUCE, a, None, None, 4, 1, lambda r, x, y, q, l: ax.EltwiseReduceMod(r, x, n, q, 4, 1)),
        (ax.OP_REDUCE, a, None, None, 4, 2, lambda r, x, y, q, l: ax.EltwiseReduceMod(r, x, n, q, 4, 2)),
        (ax.OP_REDUCE, a, None, None, ax.RNS_ANY_INPUT, 1, lambda r, x, y, q, l: ax.EltwiseReduceMod(r, x, n, q, q, 1)),
        (ax.OP_REDUCE, a, None, None, 1, 1, lambda r, x, y, q, l: ax.EltwiseReduceMod(r, x, n, q, 1, 1)),
    ]
    for i, (op, x, y, sc, fi, fo, plain) in enumerate(cases):
        out = torch.full_like(a, -1)
        ctx.eltwise(op, out, x, y, sc, count, fi, fo)
        ref = torch.full_like(a, -1)
        for p in range(count):
            for l in range(L):
                plain(ref[p, l], x[p, l], None if y is None else y[p, l], qs[l], l)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (logn, i)
    # one oracle sample: MulMod F=1 on the last limb
    out = torch.empty_like(a)
    ctx.eltwise(ax.OP_MUL, out, lo, lo_b, None, count, 1, 1)
    want = O.mul_mod(host(lo[count - 1, L - 1]), host(lo_b[count - 1, L - 1]), qs[L - 1], 1)
    assert np.array_equal(host(out[count - 1, L - 1]), want)
    # vector-scalar add against the oracle on one limb
    out = torch.empty_like(a)
    ctx.eltwise(ax.OP_ADD, out, lo, None, small, count, 1, 1)
    want = O.add_mod(host(lo[0, 1]), small[1], qs[1])
    assert np.array_equal(host(out[0, 1]), want)
    with pytest.raises(ax.AxhelError):  # add/sub need operand2 or scalars
        ctx.eltwise(ax.OP_ADD, out, lo, None, None, count, 1, 1)
    # unsupported factors leave the result untouched
    keep = torch.full_like(a, 7)
    with pytest.raises(ax.AxhelError):
        ctx.eltwise(ax.OP_MUL, keep, lo, lo_b, None, count, 3, 1)
    with pytest.raises(ax.AxhelError):
        ctx.eltwise(ax.OP_REDUCE, keep, a, None, None, count, 8, 1)
    torch.cuda.synchronize()
    assert bool((keep == 7).all())
    ctx.close()


def test_rns_calls_on_two_streams():
    """Scratch is stream-ordered: calls issued on two streams at once do not
    share job tables."""
    n, L, count = 1 << 14, 4, 16
    torch, qs, ctx, data, src = setup(n, L, count, 21, 4)
    src2 = torch.flip(src, dims=[0]).contiguous()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    o1, o2 = torch.empty_like(src), torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        ctx.ntt(o1, src, count, False, 1, stream=s1)
        ctx.ntt(o2, src2, count, False, 1, stream=s2)
    torch.cuda.synchronize()
    r1 = torch.empty_like(src)
    ctx.ntt(r1, src, count, False, 1)
    torch.cuda.synchronize()
    assert torch.equal(o1, r1)
    assert torch.equal(o2, torch.flip(r1, dims=[0]))
    ctx.close()



def test_rns_cached_tables_follow_every_change():
    """The context keeps one job table per stream and skips the fill when a
    call repeats the previous NTT call on that stream.  On one stream, in
    sequence: repeated identical calls, an element-wise call in between
    (overwrites the table), the other direction, a smaller and a larger
    count (the buffer grows), and a configuration change that alters the
    stack layout; every result equals the batch computed independently."""
    n, L = 1 << 16, 6
    torch, qs, ctx, data, src = setup(n, L, 24, 31, 4)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]

    def want(count, inverse, inp):
        out = torch.full_like(inp, -1)
        b = ax.Batch([(plans[l], out[p, l], inp[p, l], inverse, 1) for p in range(count) for l in range(L)])
        b.launch_eager()
        torch.cuda.synchronize()
        b.close()
        return out[:count]

    fwd24 = want(24, False, src)
    out = torch.full_like(src, -1)
    for _ in range(3):  # identical calls: table reused
        out.fill_(-1)
        ctx.ntt(out, src, 24, False, 1)
        torch.cuda.synchronize()
        assert torch.equal(out, fwd24)
    lo = torch.remainder(src, torch.tensor(qs, device="cuda").view(1, L, 1))
    tmp = torch.empty_like(src)
    ctx.eltwise(ax.OP_ADD, tmp, lo, lo, None, 24, 1, 1)  # overwrites the table buffer
    out.fill_(-1)
    ctx.ntt(out, src, 24, False, 1)
    torch.cuda.synchronize()
    assert torch.equal(out, fwd24)
    back = torch.full_like(src, -1)
    ctx.ntt(back, fwd24.contiguous(), 24, True, 1)  # other direction, other buffers
    torch.cuda.synchronize()
    assert torch.equal(back, torch.remainder(src, torch.tensor(qs, device="cuda").view(1, L, 1)))
    out.fill_(-1)
    ctx.ntt(out, src, 5, False, 1)  # smaller count
    torch.cuda.synchronize()
    assert torch.equal(out[:5], fwd24[:5]) and bool((out[5:] == -1).all())
    big = torch.cat([src, src], 0).contiguous()  # larger count: the buffer grows
    big_out = torch.full_like(big, -1)
    ctx.ntt(big_out, big, 48, False, 1)
    torch.cuda.synchronize()
    assert torch.equal(big_out[:24], fwd24) and torch.equal(big_out[24:], fwd24)
    try:
        for ci in range(ax.config_count(n)):  # stack layouts of every configuration
            ax.config_select(n, ci)
            out.fill_(-1)
            ctx.ntt(out, src, 24, False, 1)
            torch.cuda.synchronize()
            assert torch.equal(out, fwd24), ci
    finally:
        ax.config_select(n, -1)
    ctx.close()


def test_rns_cached_eltwise_tables_follow_every_change():
    """An element-wise call that repeats the previous call on its stream (same
    buffers, count and per-limb scalars) skips the fill; the op and factors
    are not part of the table.  In sequence on one stream: repeated calls,
    another op on the same buffers, vector-scalar calls whose scalars alone
    change, an NTT call in between, a smaller count and another result
    buffer; every result equals the per-limb calls."""
    n, L, count = 1 << 12, 5, 9
    torch, qs, ctx, a_np, a = setup(n, L, count, 41, 4)
    rng = np.random.default_rng(42)
    b = torch.from_numpy(np.stack([np.stack([params.uniform(rng, n, 4 * q) for q in qs])
                                   for _ in range(count)]).view(np.int64).copy()).cuda()
    qv = torch.tensor(qs, device="cuda").view(1, L, 1)
    lo, lo_b = torch.remainder(a, qv), torch.remainder(b, qv)
    s1 = [int(params.uniform(rng, 1, q)[0]) for q in qs]
    s2 = [int(params.uniform(rng, 1, q)[0]) for q in qs]
    s8 = [int(params.uniform(rng, 1, 8 * q)[0]) for q in qs]

    def ref(plain, x, y, cnt=count):
        out = torch.full_like(a, -1)
        for p in range(cnt):
            for l in range(L):
                plain(out[p, l], x[p, l], None if y is None else y[p, l], qs[l], l)
        torch.cuda.synchronize()
        return out

    mul = ref(lambda r, x, y, q, l: ax.EltwiseMulMod(r, x, y, n, q, 1), lo, lo_b)
    add = ref(lambda r, x, y, q, l: ax.EltwiseAddMod(r, x, y, n, q), lo, lo_b)
    add1 = ref(lambda r, x, y, q, l: ax.EltwiseAddMod(r, x, s1[l], n, q), lo, None)
    add2 = ref(lambda r, x, y, q, l: ax.EltwiseAddMod(r, x, s2[l], n, q), lo, None)
    sub2 = ref(lambda r, x, y, q, l: ax.EltwiseSubMod(r, x, s2[l], n, q), lo, None)
    fma8 = ref(lambda r, x, y, q, l: ax.EltwiseFMAMod(r, x, s8[l], y, n, q, 8), a, b)
    fma4 = ref(lambda r, x, y, q, l: ax.EltwiseFMAMod(r, x, s8[l], y, n, q, 4), lo, lo_b)
    out = torch.empty_like(a)

    def check(want, op, x, y, sc, fi, cnt=count, dst=out):
        dst.fill_(-1)
        ctx.eltwise(op, dst, x, y, sc, cnt, fi, 1)
        torch.cuda.synchronize()
        assert torch.equal(dst[:cnt], want[:cnt]) and bool((dst[cnt:] == -1).all())

    for _ in range(3):  # identical calls: table reused
        check(mul, ax.OP_MUL, lo, lo_b, None, 1)
    check(add, ax.OP_ADD, lo, lo_b, None, 1)  # same table, another op
    check(add1, ax.OP_ADD, lo, None, s1, 1)  # operand2 gone: refill
    check(add2, ax.OP_ADD, lo, None, s2, 1)  # only the scalars differ
    check(sub2, ax.OP_SUB, lo, None, s2, 1)
    check(fma8, ax.OP_FMA, a, b, s8, 8)  # scalars stored reduced for F=8
    check(fma4, ax.OP_FMA, lo, lo_b, s8, 4)  # same raw scalars, other operands and reduction
    fwd = torch.empty_like(a)
    ctx.ntt(fwd, lo, count, False, 1)  # the NTT table overwrites the buffer
    check(fma4, ax.OP_FMA, lo, lo_b, s8, 4)
    check(mul, ax.OP_MUL, lo, lo_b, None, 1, cnt=4)  # smaller count
    other = torch.empty_like(a)
    check(mul, ax.OP_MUL, lo, lo_b, None, 1, dst=other)  # another result buffer
    check(mul, ax.OP_MUL, lo, lo_b, None, 1)
    t = O.Tables(n, qs[3], params.find_primitive_2n_root(n, qs[3]))
    assert np.array_equal(host(fwd[2, 3]), O.compute_forward(host(lo[2, 3]), t, 1, 1))
    ctx.close()


def test_rns_calls_refuse_graph_capture():
    """A captured call would depend on the context's per-stream table buffer,
    which later calls overwrite: calls on a capturing stream are refused
    before anything is enqueued, and the context works normally after."""
    n, L, count = 1 << 12, 3, 2
    torch, qs, ctx, data, src = setup(n, L, count, 41, 4)
    out = torch.full_like(src, -1)
    x = torch.ones(8, device="cuda")
    y = torch.empty_like(x)
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    errors = []
    with torch.cuda.graph(g, capture_error_mode="relaxed"):
        y.copy_(x * 2)
        for call in (lambda: ctx.ntt(out, src, count, False, 1, stream=torch.cuda.current_stream()),
                     lambda: ctx.eltwise(ax.OP_ADD, out, src, src, None, count, 1, 1,
                                         stream=torch.cuda.current_stream())):
            try:
                call()
            except ax.AxhelError as e:
                errors.append(str(e))
    assert len(errors) == 2, errors
    g.replay()
    torch.cuda.synchronize()
    assert bool((out == -1).all()) and bool((y == 2).all())  # nothing of ours was captured
    ctx.ntt(out, src, count, False, 1)
    ref = torch.full_like(src, -1)
    plans = [ax.NTT(n, q, params.find_primitive_2n_root(n, q)) for q in qs]
    for p in range(count):
        for l in range(L):
            plans[l].ComputeForward(ref[p, l], src[p, l], 1, 1)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    ctx.close()
