"""ctypes front end of the CPU oracle (oracle/axhel_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it to check the HIP library.  Arrays are numpy uint64.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u64 = ctypes.c_uint64
_p = ctypes.c_void_p


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        from oracle import build_oracle
        build_oracle.build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "oracle_shoup_quotient": (_u64, [_u64, _u64]),
        "oracle_shoup_lazy": (_u64, [_u64, _u64, _u64, _u64]),
        "oracle_gen_tables": (ctypes.c_int, [_u64, _u64, _u64, _p, _p, _p]),
        "oracle_fwd_lazy_stages": (None, [_p, _u64, _u64, _p, ctypes.c_uint]),
        "oracle_fwd_lazy": (None, [_p, _u64, _u64, _p]),
        "oracle_inv_lazy_stages": (None, [_p, _u64, _u64, _p, _p, ctypes.c_uint, ctypes.c_int]),
        "oracle_inv_lazy": (None, [_p, _u64, _u64, _p, _p]),
        "oracle_add_vv": (None, [_p, _p, _p, _u64, _u64]),
        "oracle_add_vs": (None, [_p, _p, _u64, _u64, _u64]),
        "oracle_sub_vv": (None, [_p, _p, _p, _u64, _u64]),
        "oracle_sub_vs": (None, [_p, _p, _u64, _u64, _u64]),
        "oracle_mul": (None, [_p, _p, _p, _u64, _u64, _u64]),
        "oracle_fma": (None, [_p, _p, _u64, _p, _u64, _u64, _u64]),
        "oracle_reduce": (None, [_p, _p, _u64, _u64, _u64, _u64]),
        "oracle_compute_forward": (None, [_p, _p, _u64, _u64, _p, _u64, _u64]),
        "oracle_compute_inverse": (None, [_p, _p, _u64, _u64, _p, _p, _u64, _u64]),
        "oracle_is_prime": (ctypes.c_int, [_u64]),
        "oracle_find_ntt_prime": (_u64, [_u64, ctypes.c_uint]),
        "oracle_find_primitive_2n_root": (_u64, [_u64, _u64]),
        "oracle_mt_seed": (None, [_p, _u64]),
        "oracle_mt_next": (_u64, [_p]),
        "oracle_random_vector": (None, [_p, _u64, _p, _u64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous uint64"
    return a.ctypes.data


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


# ---------------------------------------------------------------- primitives
def shoup_quotient(w, q):
    return int(_load().oracle_shoup_quotient(w, q))


def shoup_lazy(v, w, wq, q):
    return int(_load().oracle_shoup_lazy(v, w, wq, q))


# ---------------------------------------------------------------- tables
class Tables:
    """Forward/inverse AoS twiddle tables ({operand, quotient} per entry)."""

    def __init__(self, n, q, psi):
        self.n, self.q, self.psi = int(n), int(q), int(psi)
        self.fwd = np.zeros(2 * n, dtype=np.uint64)
        self.inv = np.zeros(2 * n, dtype=np.uint64)
        self.ninv = np.zeros(2, dtype=np.uint64)
        rc = _load().oracle_gen_tables(n, q, psi, _ptr(self.fwd), _ptr(self.inv), _ptr(self.ninv))
        if rc != 0:
            raise MemoryError("oracle_gen_tables failed")


# ---------------------------------------------------------------- NTT
def fwd_lazy(a, tables, stop_log_m=64):
    a = _arr(a).copy()
    _load().oracle_fwd_lazy_stages(_ptr(a), tables.n, tables.q, _ptr(tables.fwd), stop_log_m)
    return a


def inv_lazy(a, tables, stages=64, final=True):
    a = _arr(a).copy()
    _load().oracle_inv_lazy_stages(_ptr(a), tables.n, tables.q, _ptr(tables.inv), _ptr(tables.ninv),
                                   stages, 1 if final else 0)
    return a


def compute_forward(a, tables, in_f=1, out_f=1):
    a = _arr(a)
    res = np.empty_like(a)
    _load().oracle_compute_forward(_ptr(res), _ptr(a), tables.n, tables.q, _ptr(tables.fwd), in_f, out_f)
    return res


def compute_inverse(a, tables, in_f=1, out_f=1):
    a = _arr(a)
    res = np.empty_like(a)
    _load().oracle_compute_inverse(_ptr(res), _ptr(a), tables.n, tables.q, _ptr(tables.inv),
                                   _ptr(tables.ninv), in_f, out_f)
    return res


def compute_forward_inplace(a, tables, out_f=1):
    """In place over a contiguous uint64 array (used for large batched checks)."""
    _load().oracle_compute_forward(_ptr(a), _ptr(a), tables.n, tables.q, _ptr(tables.fwd), 1, out_f)


def compute_inverse_inplace(a, tables, out_f=1):
    _load().oracle_compute_inverse(_ptr(a), _ptr(a), tables.n, tables.q, _ptr(tables.inv),
                                   _ptr(tables.ninv), 1, out_f)


# ---------------------------------------------------------------- element-wise
def add_mod(a, b, q):
    a = _arr(a)
    r = np.empty_like(a)
    if np.ndim(b) == 0:
        _load().oracle_add_vs(_ptr(r), _ptr(a), int(b), a.size, q)
    else:
        b = _arr(b)
        _load().oracle_add_vv(_ptr(r), _ptr(a), _ptr(b), a.size, q)
    return r


def sub_mod(a, b, q):
    a = _arr(a)
    r = np.empty_like(a)
    if np.ndim(b) == 0:
        _load().oracle_sub_vs(_ptr(r), _ptr(a), int(b), a.size, q)
    else:
        b = _arr(b)
        _load().oracle_sub_vv(_ptr(r), _ptr(a), _ptr(b), a.size, q)
    return r


def mul_mod(a, b, q, f, out=None):
    a, b = _arr(a), _arr(b)
    r = np.empty_like(a) if out is None else out
    _load().oracle_mul(_ptr(r), _ptr(a), _ptr(b), a.size, q, f)
    return r


def fma_mod(a, s, c, q, f, out=None):
    a = _arr(a)
    c = None if c is None else _arr(c)
    r = np.empty_like(a) if out is None else out
    _load().oracle_fma(_ptr(r), _ptr(a), int(s), _ptr(c), a.size, q, f)
    return r


def reduce_mod(a, q, fin, fout, out=None):
    a = _arr(a)
    r = np.empty_like(a) if out is None else out
    _load().oracle_reduce(_ptr(r), _ptr(a), a.size, q, fin, fout)
    return r


# ---------------------------------------------------------------- parameters
def is_prime(n):
    return bool(_load().oracle_is_prime(n))


def find_ntt_prime(degree, bits):
    q = int(_load().oracle_find_ntt_prime(degree, bits))
    if q == 0:
        raise RuntimeError("unable to find NTT-friendly prime")
    return q


def find_primitive_2n_root(degree, q):
    r = int(_load().oracle_find_primitive_2n_root(degree, q))
    if r == 0:
        raise RuntimeError("unable to find primitive root")
    return r


class MT19937_64:
    """mt19937_64 stream + the reference tests' rejection-sampled bounded draws."""

    def __init__(self, seed):
        self.state = np.zeros(313, dtype=np.uint64)
        _load().oracle_mt_seed(_ptr(self.state), seed & 0xFFFFFFFFFFFFFFFF)

    def next(self):
        return int(_load().oracle_mt_next(_ptr(self.state)))

    def random_vector(self, n, bound):
        out = np.empty(n, dtype=np.uint64)
        _load().oracle_random_vector(_ptr(self.state), bound, _ptr(out), n)
        return out
