"""Python host mirror of the gfx950 AXHEL library.

Thin ctypes binding of the C ABI in include/axhel_hip.h (libaxhel_hip.so),
shaped like the reference's public API (unipi::axhel, axhel/include/axhel/*.hpp):
the element-wise functions take (result, operand1, operand2/scalar, n, modulus,
factor...) and NTT(degree, modulus, root_of_unity) has ComputeForward /
ComputeInverse with the same factor conventions.  Buffers are torch uint64
(or int64) tensors on the GPU, or numpy uint64 arrays on the host (staged).

There is no CPU fallback: importing works without a GPU, but every compute
call needs the built library and a GPU, and fails loudly otherwise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libaxhel_hip.so")
CXX_LIB_PATH = os.path.join(LIB_DIR, "libaxhel.so")
CXX_CHECK_PATH = os.path.join(LIB_DIR, "axhel_cxx_check")

OK = 0
UNSUPPORTED_FACTOR = 1
INVALID_ARGUMENT = 2
NO_MEMORY = 3
HIP_ERROR = 4
NO_DEVICE = 5
REJECTED = 6
TIMEOUT = 7
CANCELLED = 8
PENDING = 9


class AxhelError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        lib = _load()
        msg = lib.axhel_hip_status_string(status).decode()
        detail = lib.axhel_hip_last_error().decode()
        super().__init__(f"{what}: {msg}" + (f" ({detail})" if detail else ""))


_lib = None
_u64 = ctypes.c_uint64
_p = ctypes.c_void_p
_st = ctypes.c_int


class NttJob(ctypes.Structure):
    """axhel_hip_ntt_job (include/axhel_hip.h)."""
    _fields_ = [("plan", _p), ("result", _p), ("operand", _p), ("output_mod_factor", _u64),
                ("inverse", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class NttConfig(ctypes.Structure):
    """axhel_hip_ntt_config (include/axhel_hip.h)."""
    _fields_ = [("log_elems", ctypes.c_int32), ("row_groups", ctypes.c_int32),
                ("col_groups", ctypes.c_int32), ("log_col", ctypes.c_int32),
                ("log_elems_col", ctypes.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class NttKernelInfo(ctypes.Structure):
    """axhel_hip_ntt_kernel_info (include/axhel_hip.h)."""
    _fields_ = [("degree", _u64), ("limbs", ctypes.c_uint32), ("pass_", ctypes.c_uint32),
                ("passes", ctypes.c_uint32), ("log_bits", ctypes.c_uint32), ("inverse", ctypes.c_int32),
                ("normalize", ctypes.c_int32), ("relaxed", ctypes.c_int32), ("stacked", ctypes.c_int32),
                ("schedule", ctypes.c_int32), ("reserved", ctypes.c_int32), ("butterflies", _u64), ("bytes", _u64)]

    def as_dict(self):
        d = {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}
        d["pass"] = d.pop("pass_")
        d["schedule_name"] = SCHEDULE_NAMES[d["schedule"]]
        return d


# butterfly schedules of the pass kernels (include/axhel_hip_ubench.h)
SCHEDULE_NAMES = ["fwd_exact", "fwd_relaxed", "fwd_relaxed_last_normalized", "inv", "inv_final",
                  "inv_final_normalized"]


class PolymulJob(ctypes.Structure):
    """axhel_hip_polymul_job (include/axhel_hip.h)."""
    _fields_ = [("plan", _p), ("result", _p), ("a", _p), ("b", _p), ("addend", _p), ("scalar", _u64),
                ("scratch", _p)]


class ServeConfig(ctypes.Structure):
    """axhel_hip_serve_config (include/axhel_hip.h)."""
    _fields_ = [("memory_budget_bytes", _u64), ("max_batch_limbs", ctypes.c_uint32), ("slots", ctypes.c_uint32),
                ("backend", ctypes.c_uint32), ("null_latency_us", ctypes.c_uint32), ("fail_every_alloc", _u64),
                ("fail_every_batch", _u64), ("latency_samples", _u64)]


class Request(ctypes.Structure):
    """axhel_hip_request (include/axhel_hip.h)."""
    _fields_ = [("op", ctypes.c_int32), ("limbs", ctypes.c_uint32), ("degree", _u64), ("plans", _p),
                ("moduli", _p), ("scalars", _p), ("result", _p), ("operand1", _p), ("operand2", _p),
                ("input_mod_factor", _u64), ("output_mod_factor", _u64), ("timeout_us", ctypes.c_int64)]


class ServeStats(ctypes.Structure):
    """axhel_hip_serve_stats (include/axhel_hip.h)."""
    _fields_ = [(k, _u64) for k in ("submitted", "completed", "rejected", "timed_out", "cancelled", "failed",
                                    "limbs_completed", "batches", "bytes_in_use", "peak_bytes_in_use")] + \
               [(k, ctypes.c_double) for k in ("p50_us", "p99_us", "mean_us", "max_us",
                                              "first_arrival_to_last_completion_s")] + \
               [("descriptor_bytes", _u64), ("pool_bytes", _u64), ("peak_pool_bytes", _u64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


OP_NTT_FORWARD, OP_NTT_INVERSE, OP_ADD, OP_SUB, OP_MUL, OP_FMA, OP_REDUCE = range(7)
BACKEND_DEVICE, BACKEND_NULL = 0, 1

# name -> (restype, argtypes); must cover every function declared in include/axhel_hip.h
SIGNATURES = {
    "axhel_hip_status_string": (ctypes.c_char_p, [_st]),
    "axhel_hip_last_error": (ctypes.c_char_p, []),
    "axhel_hip_device_count": (ctypes.c_int, []),
    "axhel_hip_is_device_pointer": (ctypes.c_int, [_p]),
    "axhel_hip_stream_synchronize": (_st, [_p]),
    "axhel_hip_add_mod_vv": (_st, [_p, _p, _p, _u64, _u64, _p]),
    "axhel_hip_add_mod_vs": (_st, [_p, _p, _u64, _u64, _u64, _p]),
    "axhel_hip_sub_mod_vv": (_st, [_p, _p, _p, _u64, _u64, _p]),
    "axhel_hip_sub_mod_vs": (_st, [_p, _p, _u64, _u64, _u64, _p]),
    "axhel_hip_mul_mod": (_st, [_p, _p, _p, _u64, _u64, _u64, _p]),
    "axhel_hip_fma_mod": (_st, [_p, _p, _u64, _p, _u64, _u64, _u64, _p]),
    "axhel_hip_reduce_mod": (_st, [_p, _p, _u64, _u64, _u64, _u64, _p]),
    "axhel_hip_ntt_plan_create": (_st, [_u64, _u64, _u64, ctypes.POINTER(_p)]),
    "axhel_hip_ntt_plan_create_external": (_st, [_u64, _u64, _p, _p, _u64, _u64, ctypes.POINTER(_p)]),
    "axhel_hip_ntt_plan_destroy": (_st, [_p]),
    "axhel_hip_ntt_plan_degree": (_u64, [_p]),
    "axhel_hip_ntt_plan_modulus": (_u64, [_p]),
    "axhel_hip_ntt_plan_tables": (_st, [_p, _p, _p, _p]),
    "axhel_hip_ntt_forward": (_st, [_p, _p, _p, _u64, _u64, _p]),
    "axhel_hip_ntt_inverse": (_st, [_p, _p, _p, _u64, _u64, _p]),
    "axhel_hip_ntt_batch_create": (_st, [ctypes.POINTER(NttJob), ctypes.c_size_t, ctypes.POINTER(_p)]),
    "axhel_hip_ntt_batch_launch": (_st, [_p, _p]),
    "axhel_hip_ntt_batch_launch_eager": (_st, [_p, _p]),
    "axhel_hip_ntt_batch_launch_graph": (_st, [_p, _p]),
    "axhel_hip_ntt_batch_destroy": (_st, [_p]),
    "axhel_hip_ntt_batch_kernel_count": (ctypes.c_int, [_p]),
    "axhel_hip_ntt_batch_launch_kernel": (_st, [_p, ctypes.c_int, _p]),
    "axhel_hip_ntt_batch_run": (_st, [ctypes.POINTER(NttJob), ctypes.c_size_t, _p]),
    "axhel_hip_ntt_batch_kernel_info": (_st, [_p, ctypes.c_int, ctypes.POINTER(NttKernelInfo)]),
    "axhel_hip_ntt_config_count": (ctypes.c_int, [_u64]),
    "axhel_hip_ntt_batch_stacked_groups": (ctypes.c_int, [_p]),
    "axhel_hip_ntt_config_get": (_st, [_u64, ctypes.c_int, ctypes.POINTER(NttConfig)]),
    "axhel_hip_ntt_config_select": (_st, [_u64, ctypes.c_int]),
    "axhel_hip_ntt_config_selected": (ctypes.c_int, [_u64]),
    "axhel_hip_ntt_config_single": (ctypes.c_int, [_u64]),
    "axhel_hip_ntt_config_small": (ctypes.c_int, [_u64]),
    "axhel_hip_polymul_create": (_st, [ctypes.POINTER(PolymulJob), ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(_p)]),
    "axhel_hip_polymul_launch": (_st, [_p, _p]),
    "axhel_hip_polymul_destroy": (_st, [_p]),
    "axhel_hip_polymul_stacked_groups": (ctypes.c_int, [_p]),
    "axhel_hip_rns_create": (_st, [_u64, _p, _p, ctypes.c_uint32, ctypes.POINTER(_p)]),
    "axhel_hip_rns_destroy": (_st, [_p]),
    "axhel_hip_rns_plan": (_p, [_p, ctypes.c_uint32]),
    "axhel_hip_rns_ntt": (_st, [_p, _p, _p, _u64, ctypes.c_int, _u64, _p]),
    "axhel_hip_rns_ntt_noalloc": (_st, [_p, _p, _p, _u64, ctypes.c_int, _u64, _p]),
    "axhel_hip_rns_eltwise": (_st, [_p, ctypes.c_int, _p, _p, _p, _p, _u64, _u64, _u64, _p]),
    "axhel_hip_server_create": (_st, [ctypes.POINTER(ServeConfig), ctypes.POINTER(_p)]),
    "axhel_hip_server_destroy": (_st, [_p]),
    "axhel_hip_server_plan": (_st, [_p, _u64, _u64, _u64, ctypes.POINTER(_p)]),
    "axhel_hip_server_submit": (_st, [_p, ctypes.POINTER(Request), ctypes.POINTER(_u64)]),
    "axhel_hip_server_wait": (_st, [_p, _u64, ctypes.c_int64]),
    "axhel_hip_server_cancel": (_st, [_p, _u64]),
    "axhel_hip_server_stats": (_st, [_p, ctypes.POINTER(ServeStats)]),
    "axhel_hip_server_reset_stats": (_st, [_p]),
    "axhel_hip_ubench_butterfly_peak": (_st, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "axhel_hip_ubench_schedule_peak": (_st, [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double)]),
}


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m axhel_amd.build` "
                          "(or __graft_entry__.build())")
    # PyTorch-ROCm ships its own libamdhip64 / libhsa-runtime64.  Loading
    # torch first lets this library bind to those (same soname), so the
    # process holds one HIP runtime; loaded the other way round, a second
    # runtime initialises the device and torch then finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    return _load()


def _check(status, what):
    if status != OK:
        raise AxhelError(status, what)


# ------------------------------------------------------------------ buffers
def _ptr(x):
    """Data pointer of a torch tensor (device or host) or numpy uint64 array."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        assert x.dtype == np.uint64 and x.flags["C_CONTIGUOUS"], "need contiguous uint64"
        return x.ctypes.data
    # torch tensor
    assert x.is_contiguous(), "need a contiguous tensor"
    assert x.element_size() == 8, "need 64-bit elements"
    return x.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream  # torch.cuda.Stream


def device_count():
    return int(_load().axhel_hip_device_count())


def synchronize(stream=None):
    _check(_load().axhel_hip_stream_synchronize(_stream_ptr(stream)), "synchronize")


# ------------------------------------------------------------ element-wise
def EltwiseAddMod(result, operand1, operand2, n, modulus, stream=None):
    """eltwise-add-mod.hpp:13,17 -- operand2 is a buffer or a Python int."""
    L = _load()
    if isinstance(operand2, (int, np.integer)):
        s = L.axhel_hip_add_mod_vs(_ptr(result), _ptr(operand1), int(operand2), n, modulus, _stream_ptr(stream))
    else:
        s = L.axhel_hip_add_mod_vv(_ptr(result), _ptr(operand1), _ptr(operand2), n, modulus, _stream_ptr(stream))
    _check(s, "EltwiseAddMod")


def EltwiseSubMod(result, operand1, operand2, n, modulus, stream=None):
    """eltwise-sub-mod.hpp:13,17."""
    L = _load()
    if isinstance(operand2, (int, np.integer)):
        s = L.axhel_hip_sub_mod_vs(_ptr(result), _ptr(operand1), int(operand2), n, modulus, _stream_ptr(stream))
    else:
        s = L.axhel_hip_sub_mod_vv(_ptr(result), _ptr(operand1), _ptr(operand2), n, modulus, _stream_ptr(stream))
    _check(s, "EltwiseSubMod")


def EltwiseMulMod(result, operand1, operand2, n, modulus, input_mod_factor, stream=None):
    """eltwise-mul-mod.hpp:14.  Returns False (output untouched) for factors the
    reference ignores, as its silent no-op."""
    s = _load().axhel_hip_mul_mod(_ptr(result), _ptr(operand1), _ptr(operand2), n, modulus, input_mod_factor,
                                  _stream_ptr(stream))
    if s == UNSUPPORTED_FACTOR:
        return False
    _check(s, "EltwiseMulMod")
    return True


def EltwiseFMAMod(res, op1, op2, op3, n, mod, mod_factor, stream=None):
    """eltwise-fma-mod.hpp:15; op3 may be None."""
    s = _load().axhel_hip_fma_mod(_ptr(res), _ptr(op1), int(op2), _ptr(op3), n, mod, mod_factor,
                                  _stream_ptr(stream))
    if s == UNSUPPORTED_FACTOR:
        return False
    _check(s, "EltwiseFMAMod")
    return True


def EltwiseReduceMod(res, op, n, mod, in_mod_factor, out_mod_factor, stream=None):
    """eltwise-reduce-mod.hpp:14."""
    s = _load().axhel_hip_reduce_mod(_ptr(res), _ptr(op), n, mod, in_mod_factor, out_mod_factor,
                                     _stream_ptr(stream))
    if s == UNSUPPORTED_FACTOR:
        return False
    _check(s, "EltwiseReduceMod")
    return True


# --------------------------------------------------------------------- NTT
class NTT:
    """ntt.hpp:24-65: NTT(degree, modulus, root_of_unity) or
    NTT(degree, modulus, root_powers=..., inv_root_powers=..., inv_degree_modulo=(op, quot)),
    tables as uint64 arrays of 2*degree words ({operand, quotient} pairs)."""

    def __init__(self, degree, modulus, root_of_unity=None, root_powers=None, inv_root_powers=None,
                 inv_degree_modulo=None):
        L = _load()
        h = _p()
        if root_of_unity is not None:
            s = L.axhel_hip_ntt_plan_create(degree, modulus, root_of_unity, ctypes.byref(h))
        else:
            # borrowed for the plan's life (ntt.hpp:34-37): a uint64 contiguous
            # array is used in place, so later changes to it reach the calls
            rp = np.ascontiguousarray(root_powers, dtype=np.uint64)
            ip = np.ascontiguousarray(inv_root_powers, dtype=np.uint64)
            self._tables = (rp, ip)
            s = L.axhel_hip_ntt_plan_create_external(degree, modulus, rp.ctypes.data, ip.ctypes.data,
                                                     int(inv_degree_modulo[0]), int(inv_degree_modulo[1]),
                                                     ctypes.byref(h))
        _check(s, "NTT")
        self._h = h
        self.degree = int(degree)
        self.modulus = int(modulus)

    @property
    def handle(self):
        return self._h.value

    def Degree(self):
        return int(_load().axhel_hip_ntt_plan_degree(self._h))

    def Modulus(self):
        return int(_load().axhel_hip_ntt_plan_modulus(self._h))

    def tables(self):
        n = self.degree
        fwd = np.zeros(2 * n, dtype=np.uint64)
        inv = np.zeros(2 * n, dtype=np.uint64)
        ni = np.zeros(2, dtype=np.uint64)
        _check(_load().axhel_hip_ntt_plan_tables(self._h, fwd.ctypes.data, inv.ctypes.data, ni.ctypes.data),
               "tables")
        return fwd, inv, ni

    def ComputeForward(self, result, operand, input_mod_factor=1, output_mod_factor=1, stream=None):
        _check(_load().axhel_hip_ntt_forward(self._h, _ptr(result), _ptr(operand), input_mod_factor,
                                             output_mod_factor, _stream_ptr(stream)), "ComputeForward")

    def ComputeInverse(self, result, operand, input_mod_factor=1, output_mod_factor=1, stream=None):
        _check(_load().axhel_hip_ntt_inverse(self._h, _ptr(result), _ptr(operand), input_mod_factor,
                                             output_mod_factor, _stream_ptr(stream)), "ComputeInverse")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _load().axhel_hip_ntt_plan_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- batching
class Batch:
    """A reusable batch of limb-transforms captured into one hipGraph.

    jobs: iterable of (ntt, result_tensor, operand_tensor, inverse: bool,
    output_mod_factor) with device tensors of ntt.degree words each (views into
    one big tensor are fine)."""

    def __init__(self, jobs):
        jobs = list(jobs)
        arr = (NttJob * max(1, len(jobs)))()
        self._keep = []
        for i, (ntt, res, op, inverse, out_f) in enumerate(jobs):
            arr[i].plan = ntt.handle
            arr[i].result = res if isinstance(res, int) else _ptr(res)
            arr[i].operand = op if isinstance(op, int) else _ptr(op)
            arr[i].output_mod_factor = out_f
            arr[i].inverse = 1 if inverse else 0
            self._keep.append(ntt)
        h = _p()
        _check(_load().axhel_hip_ntt_batch_create(arr, len(jobs), ctypes.byref(h)), "batch_create")
        self._h = h

    def launch(self, stream=None):
        """One (degree, direction) group: direct kernel launches; several: the graph."""
        _check(_load().axhel_hip_ntt_batch_launch(self._h, _stream_ptr(stream)), "batch_launch")

    def launch_graph(self, stream=None):
        _check(_load().axhel_hip_ntt_batch_launch_graph(self._h, _stream_ptr(stream)), "batch_launch_graph")

    def launch_eager(self, stream=None):
        _check(_load().axhel_hip_ntt_batch_launch_eager(self._h, _stream_ptr(stream)), "batch_launch_eager")

    def kernel_count(self):
        return int(_load().axhel_hip_ntt_batch_kernel_count(self._h))

    def launch_kernel(self, index, stream=None):
        _check(_load().axhel_hip_ntt_batch_launch_kernel(self._h, index, _stream_ptr(stream)), "launch_kernel")

    def kernel_info(self, index):
        """What the index-th launch of a replay computes: degree, limbs,
        pass, butterflies, algorithmic bytes and butterfly schedule."""
        info = NttKernelInfo()
        _check(_load().axhel_hip_ntt_batch_kernel_info(self._h, index, ctypes.byref(info)), "kernel_info")
        return info.as_dict()

    def stacked_groups(self):
        """Groups whose ROW pass runs over stacks of same-plan jobs."""
        return int(_load().axhel_hip_ntt_batch_stacked_groups(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _load().axhel_hip_ntt_batch_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def butterfly_peak(schedule=0):
    """Device-wide rate of register-resident butterflies of one pass
    schedule (SCHEDULE_NAMES; 0 = the exact forward butterfly), and the
    timed launch's duration in seconds."""
    r, s = ctypes.c_double(), ctypes.c_double()
    _check(_load().axhel_hip_ubench_schedule_peak(int(schedule), ctypes.byref(r), ctypes.byref(s)),
           "butterfly_peak")
    return r.value, s.value


# ------------------------------------------------------- polynomial product
RNS_ANY_INPUT = (1 << 64) - 1


class RNS:
    """RNS context (axhel_hip_rns_*): degree N, L moduli (roots optional),
    operating on device tensors of count x L x N words."""

    def __init__(self, degree, moduli, roots=None):
        moduli = [int(q) for q in moduli]
        qa = (_u64 * len(moduli))(*moduli)
        ra = None if roots is None else (_u64 * len(moduli))(*[int(r) for r in roots])
        h = _p()
        _check(_load().axhel_hip_rns_create(degree, qa, ra, len(moduli), ctypes.byref(h)), "rns_create")
        self._h = h
        self.degree, self.moduli = degree, moduli

    def ntt(self, result, operand, count, inverse=False, output_mod_factor=1, stream=None):
        _check(_load().axhel_hip_rns_ntt(self._h, _ptr(result), _ptr(operand), count, 1 if inverse else 0,
                                         output_mod_factor, _stream_ptr(stream)), "rns_ntt")

    def ntt_noalloc(self, result, operand, count, inverse=False, output_mod_factor=1, stream=None):
        """ntt() through the entry point that allocates nothing (tables in the
        library's fixed per-GPU scratch array; calls overlapping on one GPU race)."""
        _check(_load().axhel_hip_rns_ntt_noalloc(self._h, _ptr(result), _ptr(operand), count, 1 if inverse else 0,
                                                 output_mod_factor, _stream_ptr(stream)), "rns_ntt_noalloc")

    def eltwise(self, op, result, operand1, operand2=None, scalars=None, count=1, input_mod_factor=1,
                output_mod_factor=1, stream=None):
        sa = None if scalars is None else (_u64 * len(self.moduli))(*[int(x) for x in scalars])
        _check(_load().axhel_hip_rns_eltwise(self._h, op, _ptr(result), _ptr(operand1),
                                             None if operand2 is None else _ptr(operand2), sa, count,
                                             input_mod_factor, output_mod_factor, _stream_ptr(stream)),
               "rns_eltwise")

    def plan_handle(self, limb):
        return _load().axhel_hip_rns_plan(self._h, limb)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _load().axhel_hip_rns_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Polymul:
    """Batched RNS polynomial product: per job (ntt, result, a, b, addend or
    None, scalar, scratch) computes InverseNTT(FMA(Mul(NTT(a), NTT(b)), scalar,
    addend)) with device tensors of ntt.degree words."""

    def __init__(self, jobs, fused=True):
        jobs = list(jobs)
        arr = (PolymulJob * max(1, len(jobs)))()
        self._keep = []
        for i, (ntt, res, a, b, c, s, scr) in enumerate(jobs):
            arr[i].plan = ntt.handle
            arr[i].result = _ptr(res)
            arr[i].a = _ptr(a)
            arr[i].b = _ptr(b)
            arr[i].addend = None if c is None else _ptr(c)
            arr[i].scalar = s
            arr[i].scratch = _ptr(scr)
            self._keep.append(ntt)
        h = _p()
        _check(_load().axhel_hip_polymul_create(arr, len(jobs), 1 if fused else 0, ctypes.byref(h)), "polymul_create")
        self._h = h

    def launch(self, stream=None):
        _check(_load().axhel_hip_polymul_launch(self._h, _stream_ptr(stream)), "polymul_launch")

    def stacked_groups(self):
        """Degree groups whose fused kernel runs over stacks of same-plan jobs."""
        return int(_load().axhel_hip_polymul_stacked_groups(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _load().axhel_hip_polymul_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ serving
class Server:
    """Continuous-batching request loop (axhel_hip_server_*).

    submit() takes one request over `limbs` vectors of `degree` words (limb l
    at offset l*degree of every buffer) and returns a ticket; wait(ticket)
    returns the final status code (OK, TIMEOUT, CANCELLED, NO_MEMORY, ...),
    or PENDING when wait's own timeout passed first (the ticket stays valid)."""

    def __init__(self, memory_budget_bytes=0, max_batch_limbs=0, slots=0, backend=BACKEND_DEVICE,
                 null_latency_us=0, fail_every_alloc=0, fail_every_batch=0, latency_samples=0):
        cfg = ServeConfig(memory_budget_bytes, max_batch_limbs, slots, backend, null_latency_us, fail_every_alloc,
                          fail_every_batch, latency_samples)
        h = _p()
        _check(_load().axhel_hip_server_create(ctypes.byref(cfg), ctypes.byref(h)), "server_create")
        self._h = h

    def plan(self, degree, modulus, root):
        p = _p()
        _check(_load().axhel_hip_server_plan(self._h, degree, modulus, root, ctypes.byref(p)), "server_plan")
        return p.value

    def submit(self, op, degree, limbs, result, operand1, operand2=None, plans=None, moduli=None, scalars=None,
               input_mod_factor=1, output_mod_factor=1, timeout_us=0):
        """Returns (status, ticket); ticket is None unless status == OK."""
        keep = []

        def arr(vals):
            if vals is None:
                return None
            a = np.ascontiguousarray(np.asarray(vals, dtype=np.uint64))
            keep.append(a)
            return a.ctypes.data

        r = Request()
        r.op, r.limbs, r.degree = op, limbs, degree
        if plans is not None:
            pa = (_p * limbs)(*[p if isinstance(p, int) else p.handle for p in plans])
            keep.append(pa)
            r.plans = ctypes.cast(pa, _p)
        r.moduli = arr(moduli)
        r.scalars = arr(scalars)
        r.result = result if isinstance(result, int) else _ptr(result)
        r.operand1 = operand1 if isinstance(operand1, int) else _ptr(operand1)
        r.operand2 = operand2 if (operand2 is None or isinstance(operand2, int)) else _ptr(operand2)
        r.input_mod_factor, r.output_mod_factor, r.timeout_us = input_mod_factor, output_mod_factor, timeout_us
        t = _u64()
        s = _load().axhel_hip_server_submit(self._h, ctypes.byref(r), ctypes.byref(t))
        return s, (t.value if s == OK else None)

    def wait(self, ticket, timeout_us=-1):
        return int(_load().axhel_hip_server_wait(self._h, ticket, timeout_us))

    def cancel(self, ticket):
        return int(_load().axhel_hip_server_cancel(self._h, ticket))

    def stats(self):
        st = ServeStats()
        _check(_load().axhel_hip_server_stats(self._h, ctypes.byref(st)), "server_stats")
        return st.as_dict()

    def reset_stats(self):
        _check(_load().axhel_hip_server_reset_stats(self._h), "server_reset_stats")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _load().axhel_hip_server_destroy(self._h)
            self._h = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ tuning
def config_count(degree):
    return int(_load().axhel_hip_ntt_config_count(degree))


def config_get(degree, index):
    c = NttConfig()
    _check(_load().axhel_hip_ntt_config_get(degree, index, ctypes.byref(c)), "config_get")
    return c.as_dict()


def config_select(degree, index):
    _check(_load().axhel_hip_ntt_config_select(degree, index), "config_select")


def config_selected(degree):
    return int(_load().axhel_hip_ntt_config_selected(degree))


def config_single(degree):
    """Index of the configuration single-limb calls use for this degree."""
    return int(_load().axhel_hip_ntt_config_single(degree))


def config_small(degree):
    """Index of the configuration launches of 2 limbs up to 2^21 words use."""
    return int(_load().axhel_hip_ntt_config_small(degree))
