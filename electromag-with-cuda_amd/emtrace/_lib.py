"""ctypes binding of libemtrace.so (the C ABI declared in include/emtrace.h).

The library is loaded from ``electromag-with-cuda_amd/lib/libemtrace.so`` (built
in-tree by ``make -C electromag-with-cuda_amd``).  There is no fallback: if the
library is missing every entry point raises ``LibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libemtrace.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "emtrace.h")

# status codes (include/emtrace.h)
EM_OK = 0
EM_ERR_NO_LINES = 1
EM_ERR_RESOLUTION = 2
EM_ERR_LENGTH = 3
EM_ERR_INVALID = 16
EM_ERR_HIP = 17
EM_ERR_NOMEM = 18
EM_ERR_NO_DEVICE = 19
EM_ERR_UNSUPPORTED = 20


class LibraryMissing(RuntimeError):
    pass


class EmError(RuntimeError):
    def __init__(self, code: int, message: str, func: str = ""):
        super().__init__(f"{func}: status {code}: {message}")
        self.code = code


class em_step(ctypes.Structure):
    _fields_ = [("rule", ctypes.c_int32), ("resolution", ctypes.c_float),
                ("h0", ctypes.c_float), ("hmin", ctypes.c_float),
                ("hmax", ctypes.c_float), ("turn", ctypes.c_float), ("stop", ctypes.c_int32)]


class em_launch(ctypes.Structure):
    _fields_ = [("lanes_per_line", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("block", ctypes.c_int32), ("unroll", ctypes.c_int32)]


class em_perf(ctypes.Structure):
    _fields_ = [("time_s", ctypes.c_double), ("kernel_s", ctypes.c_double),
                ("h2d_s", ctypes.c_double), ("d2h_s", ctypes.c_double),
                ("h2d_bytes", ctypes.c_double), ("d2h_bytes", ctypes.c_double),
                ("interactions", ctypes.c_double), ("interactions_per_s", ctypes.c_double),
                ("line_steps_per_s", ctypes.c_double), ("gflops", ctypes.c_double),
                ("progress", ctypes.c_double), ("lanes_per_line", ctypes.c_int32),
                ("tile", ctypes.c_int32), ("block", ctypes.c_int32), ("unroll", ctypes.c_int32),
                ("progress_out", ctypes.POINTER(ctypes.c_double))]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "progress_out"}


class em_batch_cfg(ctypes.Structure):
    _fields_ = [("charge_offset", ctypes.c_uint64), ("line_offset", ctypes.c_uint64),
                ("pitch", ctypes.c_uint64), ("p", ctypes.c_uint32), ("n", ctypes.c_uint32),
                ("len", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


P = ctypes.c_void_p
SZ = ctypes.c_size_t
I = ctypes.c_int
U32 = ctypes.c_uint32
F = ctypes.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    "em_init": (I, [I]),
    "em_shutdown": (I, []),
    "em_last_error": (ctypes.c_char_p, []),
    "em_device_count": (I, []),
    "em_version": (ctypes.c_char_p, []),
    "em_calc_lines_f32": (I, [P, SZ, P, P, P, SZ, SZ, F, I, ctypes.POINTER(em_perf)]),
    "em_trace_host_f32": (I, [P, SZ, P, P, P, SZ, SZ, ctypes.POINTER(em_step),
                              ctypes.POINTER(em_launch), ctypes.POINTER(em_perf), P]),
    "em_trace_dev_f32": (I, [P, SZ, P, P, P, SZ, SZ, SZ, ctypes.POINTER(em_step),
                             ctypes.POINTER(em_launch), P, U32, P, P]),
    "em_trace_dev_f32_noalloc": (I, [P, SZ, P, P, P, SZ, SZ, SZ, ctypes.POINTER(em_step),
                                     ctypes.POINTER(em_launch), P, U32, P, P]),
    "em_ckpt_floats": (SZ, [SZ, SZ, U32]),
    "em_field_dev_f32": (I, [P, SZ, P, P, P, SZ, F, P, P, P, ctypes.POINTER(em_launch), P]),
    "em_batch_create": (I, [ctypes.POINTER(em_batch_cfg), SZ, ctypes.POINTER(em_step),
                            ctypes.POINTER(em_launch), I, ctypes.POINTER(P)]),
    "em_batch_run": (I, [P, P, P, P, P, P, P]),
    "em_batch_info": (I, [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                          ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(em_launch)]),
    "em_batch_destroy": (I, [P]),
    "em_adjoint_dev_f32": (I, [P, SZ, P, P, P, SZ, SZ, SZ, ctypes.POINTER(em_step),
                               ctypes.POINTER(em_launch), P, U32, P, P, P, P, P, P]),
    "em_batch_ckpt_floats": (SZ, [P, U32]),
    "em_batch_run_ckpt": (I, [P, P, P, P, P, P, U32, P, P]),
    "em_batch_adjoint": (I, [P, P, P, P, P, P, U32, P, P, P, P, P]),
    "em_calc_lines_f64": (I, [P, SZ, P, P, P, SZ, SZ, ctypes.c_double, I, ctypes.POINTER(em_perf)]),
    "em_trace_host_f64": (I, [P, SZ, P, P, P, SZ, SZ, ctypes.POINTER(em_step),
                              ctypes.POINTER(em_perf), P]),
    "em_trace_dev_f64": (I, [P, SZ, P, P, P, SZ, SZ, SZ, ctypes.POINTER(em_step), P, P]),
    "em_verlet_step_f32": (I, [P, SZ, P, SZ, P, P, P, F, I, ctypes.POINTER(em_perf)]),
    "em_verlet_step_dev_f32": (I, [P, SZ, P, SZ, P, P, P, F, P, P]),
    "em_legacy_rand": (I, [U32, P, SZ]),
    "em_legacy_problem_f32": (I, [P, SZ, U32, P, P, P, SZ, SZ, SZ, I, U32]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libemtrace.so (once).  Raises LibraryMissing if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} not found: build it with `make -C {PKG_ROOT}` "
                "(or __graft_entry__.build()); the tracer has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, func: str = "") -> None:
    if rc != EM_OK:
        msg = lib().em_last_error()
        raise EmError(rc, msg.decode() if msg else "", func)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
