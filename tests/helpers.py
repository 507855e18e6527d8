"""Shared helpers for the GPU parity tests: device buffers and reference
parameter sets (tests/common/parameters.hpp:158-251 restated)."""
import numpy as np

from axhel_amd import params

MODULUS_BITS = [20, 30, 40, 50, 52, 55, 59, 60, 61, 62]  # parameters.hpp:200-205
NTT_DEGREES = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768]  # :246-251
SEED = 0x415848454C  # parameters.hpp:142-146


def element_sizes(lanes=2):
    """parameters.hpp:158-183 (with the surrogate lane count of Native builds)."""
    v = [0, 1, 2, 3, 4, 5, 6, 7, 31, 32, 33, 127, 128, 129, 255, 256, 257, 1023, 1024, 1025,
         8191, 8192, 8193, 16383, 16384, 16385, 32767, 32768, 32769]
    for k in (1, 2, 3, 4, 7, 8, 16):
        x = k * lanes
        v += [x - 1, x, x + 1]
    return sorted(set(v))


def boundary_values(q, factor):
    """parameters.hpp:272-293."""
    B = q * factor
    v = {0, 1}
    for k in range(1, factor):
        t = k * q
        for x in (t - 2, t - 1, t, t + 1):
            if 0 <= x < B and x < 1 << 64:
                v.add(x)
    for x in (B - 2, B - 1):
        if 0 <= x < B and x < 1 << 64:
            v.add(x)
    return sorted(v)


def make_pattern(n, vals):
    return np.array([vals[i % len(vals)] for i in range(n)], dtype=np.uint64)


def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def dev(a):
    """numpy uint64 -> contiguous int64-typed device tensor with the same bits."""
    torch = torch_mod()
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64).copy()).to("cuda")


def empty_dev(n):
    torch = torch_mod()
    return torch.empty(max(n, 0), dtype=torch.int64, device="cuda")


def host(t):
    torch = torch_mod()
    torch.cuda.synchronize()
    return t.detach().to("cpu").numpy().view(np.uint64).copy()


def ntt_params(n, bits):
    q = params.find_ntt_prime(n, bits)
    return q, params.find_primitive_2n_root(n, q)
