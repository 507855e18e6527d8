"""Workload parameters: NTT-friendly primes, primitive roots, residue vectors.

Python restatement of the reference test/benchmark generators
(tests/common/parameters.hpp:29-122): deterministic Miller-Rabin, the largest
prime q < 2^bits with q == 1 (mod 2N), and the smallest-base primitive 2N-th
root.  Residue vectors come from numpy's PCG64 (uniform in [0, bound)); they
are inputs, not part of the arithmetic.
"""
import numpy as np

_SMALL = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
_BASES = (2, 325, 9375, 28178, 450775, 9780504, 1795265022)


def is_prime(n):
    """parameters.hpp:29-51."""
    if n < 2:
        return False
    for p in _SMALL:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _BASES:
        if a % n == 0:
            continue
        x = pow(a % n, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(1, s):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def ntt_primes(degree, bits, count=1):
    """`count` distinct primes q < 2^bits, q == 1 mod 2*degree, walking down
    from the largest (parameters.hpp:58-70 for count == 1)."""
    step = 2 * degree
    hi = (1 << bits) - 1
    lo = 1 << (bits - 1)
    q = hi - ((hi - 1) % step)
    out = []
    while q >= lo and q > step and len(out) < count:
        if is_prime(q):
            out.append(q)
        q -= step
    if len(out) < count:
        raise RuntimeError(f"only {len(out)} NTT primes of {bits} bits for N={degree}")
    return out


def find_ntt_prime(degree, bits):
    return ntt_primes(degree, bits, 1)[0]


def find_primitive_2n_root(degree, q):
    """parameters.hpp:110-122."""
    order = 2 * degree
    e = (q - 1) // order
    for a in range(2, 100000):
        psi = pow(a, e, q)
        if psi != 1 and pow(psi, degree, q) == q - 1 and pow(psi, order, q) == 1:
            return psi
    raise RuntimeError("unable to find primitive root")


def uniform(rng, shape, bound):
    """uint64 residues uniform in [0, bound) (bound <= 2^64)."""
    if bound >= (1 << 64):
        return rng.integers(0, 1 << 64, size=shape, dtype=np.uint64, endpoint=False)
    return rng.integers(0, bound, size=shape, dtype=np.uint64)


def rns_primes(degree, limbs, bit_schedule=range(50, 61)):
    """`limbs` distinct NTT primes with bit widths cycling through
    bit_schedule, each the next unused prime of its width from the top."""
    bits = list(bit_schedule)
    need = {}
    for i in range(limbs):
        b = bits[i % len(bits)]
        need[b] = need.get(b, 0) + 1
    pools = {b: ntt_primes(degree, b, c) for b, c in need.items()}
    out = []
    used = {b: 0 for b in pools}
    for i in range(limbs):
        b = bits[i % len(bits)]
        out.append(pools[b][used[b]])
        used[b] += 1
    return out
