"""The no-allocation RNS transform (axhel_hip_rns_ntt_noalloc) gives the same
bits as axhel_hip_rns_ntt, called through the C ABI: forward and inverse,
lazy and normalized, two calls in a row with the second on a smaller input
(the fixed scratch array keeps the first call's tables, which the second
call must overwrite), single-call and stacked batch configurations, and a
call beyond the scratch maximum is refused without writing."""
import ctypes

import numpy as np
import pytest

import axhel_amd as ax
from axhel_amd import params
from tests.helpers import torch_mod

pytestmark = pytest.mark.gpu


def setup(n, L, count, seed):
    torch = torch_mod()
    qs = params.rns_primes(n, L)
    ctx = ax.RNS(n, qs)
    rng = np.random.default_rng(seed)
    data = np.stack([np.stack([params.uniform(rng, n, 4 * q) for q in qs]) for _ in range(count)])
    return torch, ctx, torch.from_numpy(data.view(np.int64).copy()).cuda()


def noalloc(ctx, out, src, count, inverse, factor):
    """The entry point called directly through ctypes."""
    return ax.lib().axhel_hip_rns_ntt_noalloc(ctx._h, out.data_ptr(), src.data_ptr(), count, 1 if inverse else 0,
                                              factor, None)


@pytest.mark.parametrize("logn,L,count,small", [(12, 4, 24, 5), (16, 24, 64, 3)])
def test_noalloc_equals_the_allocating_call(logn, L, count, small):
    n = 1 << logn
    print(f"noalloc case N=2^{logn} L={L} count={count} then {small}", flush=True)
    torch, ctx, src = setup(n, L, count, logn + L)
    for inverse in (False, True):
        for factor in ((2, 1) if inverse else (4, 1)):
            for cnt in (count, small):  # two calls in a row, the second smaller
                want = torch.full_like(src, -1)
                ctx.ntt(want, src, cnt, inverse, factor)
                got = torch.full_like(src, -1)
                assert noalloc(ctx, got, src, cnt, inverse, factor) == ax.OK
                torch.cuda.synchronize()
                assert torch.equal(got[:cnt], want[:cnt]), (logn, inverse, factor, cnt)
                assert bool((got[cnt:] == -1).all())  # nothing written past the call's polynomials
            # in place
            buf = src[:small].clone()
            assert noalloc(ctx, buf, buf, small, inverse, factor) == ax.OK
            want = torch.full_like(buf, -1)
            ctx.ntt(want, src[:small], small, inverse, factor)
            torch.cuda.synchronize()
            assert torch.equal(buf, want)
    ctx.close()


def test_noalloc_refuses_calls_beyond_the_scratch_maximum():
    torch = torch_mod()
    n, L = 1 << 10, 8
    qs = params.rns_primes(n, L)
    ctx = ax.RNS(n, qs)
    count = ax_max = 16384 // L + 1  # more descriptors than the scratch holds
    src = torch.zeros((count, L, n), dtype=torch.int64, device="cuda")
    out = torch.full_like(src, -7)
    st = noalloc(ctx, out, src, count, False, 1)
    assert st == ax.INVALID_ARGUMENT, ax_max
    assert "scratch maximum" in ax.lib().axhel_hip_last_error().decode()
    torch.cuda.synchronize()
    assert bool((out == -7).all())
    ctx.close()
