"""The oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker (or the timed CPU
baseline), never as the thing measured or shipped.  The product library
(``electromag-with-cuda_amd/``) does not import it.

Contents
--------
* ``restate.c``  -- C restatement of the reference's CPU tracer
  (ElectroMag/src/CPU_Implement.cpp:38-102, 207-351, 366-498) plus the
  project-defined angle rule in the same op order as the HIP step code.
* ``oracle64.c`` -- FP64 field, field-gradient JVP, trajectory and adjoint.
* ``ref_driver.cpp`` + ``ref_init_driver.cpp`` + ``Makefile ref`` -- the
  reference's own CPU_Implement.cpp and the problem initialisation of its
  Electromag utils.h compiled from /root/reference into ``_ref/libemref.so``
  (absent where the reference is absent; tests then use the golden vectors
  under tests/golden/, which that build wrote).

Every function below cites the reference code it follows; the pins that check
each one are listed in tests/test_oracle.py.  Parity status: all functions are
pinned (golden vectors from the compiled reference, closed forms, invariants,
finite differences); none is "parity unpinned".
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# EMTRACE_ORACLE_LIB: another build of the same sources (the sanitizer build)
_LIB_PATH = os.environ.get("EMTRACE_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(HERE, "_ref", "libemref.so")
REFERENCE_DIR = "/root/reference"

PLAIN, REF, ANGLE = 0, 1, 2
K_E = 8.987551787e9


class _StepParams(ctypes.Structure):
    _fields_ = [("rule", ctypes.c_int), ("resolution", ctypes.c_double),
                ("h0", ctypes.c_double), ("hmin", ctypes.c_double),
                ("hmax", ctypes.c_double), ("turn", ctypes.c_double), ("stop", ctypes.c_int)]


def build(ref: bool = True) -> None:
    """Compile the C oracle (and the reference checker when the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir(REFERENCE_DIR):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build(ref=False)
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        L.orc_calc_lines_f32.argtypes = [P, S, P, P, P, S, S, ctypes.POINTER(_StepParams), P]
        L.orc_calc_lines_f64.argtypes = [P, S, P, P, P, S, S, ctypes.POINTER(_StepParams), P]
        L.orc_field_f32.argtypes = [P, S, P, S, ctypes.c_float, P]
        L.orc_field_f64.argtypes = [P, S, P, S, ctypes.c_double, P]
        L.o64_field.argtypes = [P, S, P, S, ctypes.c_double, P]
        L.o64_field_jvp.argtypes = [P, S, P, S, ctypes.c_double, P, P]
        L.o64_trace.argtypes = [P, S, P, S, S, ctypes.POINTER(_StepParams), ctypes.c_double, P, P]
        L.o64_trace_jitter.argtypes = [P, S, P, S, S, ctypes.POINTER(_StepParams), ctypes.c_double,
                                       P, P, ctypes.c_double, ctypes.c_ulonglong]
        L.o64_trace_jitter.restype = ctypes.c_int
        L.o64_adjoint.argtypes = [P, S, P, S, S, ctypes.POINTER(_StepParams),
                                  ctypes.c_double, P, P, P]
        L.o64_adjoint_path.argtypes = [P, S, P, S, S, ctypes.POINTER(_StepParams),
                                       ctypes.c_double, P, P, P, P, P]
        L.o64_adjoint_path.restype = ctypes.c_int
        L.o64_verlet_step.argtypes = [P, S, P, S, P, P, P, ctypes.c_double, P]
        L.o64_verlet_step.restype = None
        L.o64_min_dist_f32.argtypes = [P, S, P, S, P]
        L.o64_path_dist_f32.argtypes = [P, P, S, S, P, ctypes.c_longlong, P]
        for f in ("orc_calc_lines_f32", "orc_calc_lines_f64", "o64_trace", "o64_adjoint"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def ref_lib() -> Optional[ctypes.CDLL]:
    """The compiled reference (oracle/_ref), or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        L = ctypes.CDLL(_REF_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        L.ref_calc_lines_f32.argtypes = [P, S, P, P, P, S, S, ctypes.c_float, ctypes.c_int, P, P]
        L.ref_calc_lines_f64.argtypes = [P, S, P, P, P, S, S, ctypes.c_double, ctypes.c_int, P, P]
        L.ref_calc_lines_f32.restype = ctypes.c_int
        L.ref_calc_lines_f64.restype = ctypes.c_int
        L.ref_init_charges_f32.argtypes = [P, S, ctypes.c_int]
        L.ref_init_charges_f32.restype = None
        L.ref_init_lines_f32.argtypes = [P, P, P, S, S, S]
        L.ref_init_lines_f32.restype = None
        _ref = L
    return _ref


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _params(rule) -> _StepParams:
    return _StepParams(int(rule.rule), float(rule.resolution), float(rule.h0),
                       float(rule.hmin), float(rule.hmax), float(rule.turn), int(rule.stop))


def _rows_from_seeds(seeds: np.ndarray, length: int, dtype) -> np.ndarray:
    n = seeds.shape[0]
    rows = np.zeros((3, length, n), dtype=dtype)
    rows[:, 0, :] = seeds.T
    return rows


def calc_lines(charges, seeds, length, rule, dtype=np.float32, with_steps=False):
    """Restated reference tracer.  Returns (rc, rows[3, len, n]) [, steps[n]]."""
    charges = np.ascontiguousarray(charges, dtype=dtype)
    rows = _rows_from_seeds(np.asarray(seeds), length, dtype)
    steps = np.zeros(rows.shape[2], dtype=np.uint32)
    f = lib().orc_calc_lines_f32 if dtype == np.float32 else lib().orc_calc_lines_f64
    sp = _params(rule)
    rc = f(_ptr(charges), charges.shape[0], _ptr(rows[0]), _ptr(rows[1]), _ptr(rows[2]),
           rows.shape[2], length, ctypes.byref(sp), _ptr(steps))
    return (rc, rows, steps) if with_steps else (rc, rows)


def ref_calc_lines(charges, seeds, length, resolution, use_curvature, dtype=np.float32):
    """The reference's own CalcField_CPU<T>.  Returns (rc, rows, time_s, gflops)."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libemref.so not built (reference absent)")
    charges = np.ascontiguousarray(charges, dtype=dtype)
    rows = _rows_from_seeds(np.asarray(seeds), length, dtype)
    t = ctypes.c_double(0)
    g = ctypes.c_double(0)
    f = L.ref_calc_lines_f32 if dtype == np.float32 else L.ref_calc_lines_f64
    rc = f(_ptr(charges), charges.shape[0], _ptr(rows[0]), _ptr(rows[1]), _ptr(rows[2]),
           rows.shape[2], length, resolution, int(use_curvature), ctypes.byref(t), ctypes.byref(g))
    return rc, rows, t.value, g.value


def ref_init_charges(p: int) -> np.ndarray:
    """The reference's InitializePointChargeArray<float>(charges, p, false)
    (ElectroMag/src/Electromag utils.h:91-111: srand(1), four rand() draws per
    charge).  Returns (p, 4) float32 {x, y, z, q}."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libemref.so not built (reference absent)")
    out = np.empty((p, 4), np.float32)
    L.ref_init_charges_f32(_ptr(out), p, 0)
    return out


def ref_init_grid(width: int, height: int, depth: int) -> np.ndarray:
    """The reference's InitializeFieldLineArray<float>(..., random=false)
    (Electromag utils.h:29-70).  Returns (n, 3) float32 seeds."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libemref.so not built (reference absent)")
    n = width * height * depth
    xyz = np.empty((3, n), np.float32)
    L.ref_init_lines_f32(_ptr(xyz[0]), _ptr(xyz[1]), _ptr(xyz[2]), width, height, depth)
    return np.ascontiguousarray(xyz.T)


def field_ref_f32(charges, pts, k=np.float32(K_E)):
    charges = np.ascontiguousarray(charges, dtype=np.float32)
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    E = np.zeros_like(pts)
    lib().orc_field_f32(_ptr(charges), charges.shape[0], _ptr(pts), pts.shape[0], k, _ptr(E))
    return E


def field64(charges, pts, scale=1.0):
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    E = np.zeros_like(pts)
    lib().o64_field(_ptr(charges), charges.shape[0], _ptr(pts), pts.shape[0], scale, _ptr(E))
    return E


def field_jvp64(charges, pts, w, scale=1.0):
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros_like(pts)
    lib().o64_field_jvp(_ptr(charges), charges.shape[0], _ptr(pts), pts.shape[0], scale,
                        _ptr(w), _ptr(out))
    return out


def min_dist(charges, rows):
    """Distance of every point of rows[3, len, n] to the nearest charge -> (len, n)."""
    charges = np.ascontiguousarray(charges, dtype=np.float32)
    pts = np.ascontiguousarray(np.asarray(rows, dtype=np.float32).transpose(1, 2, 0))
    out = np.zeros(pts.shape[:2], dtype=np.float32)
    lib().o64_min_dist_f32(_ptr(charges), charges.shape[0], _ptr(pts), pts.shape[0] * pts.shape[1],
                           _ptr(out))
    return out


def path_dist(test_rows, ref_rows, first, window=64):
    """Distance of each test point to the ref polyline of its line -> (len, n)."""
    t = np.ascontiguousarray(test_rows, dtype=np.float32)
    r = np.ascontiguousarray(ref_rows, dtype=np.float32)
    f = np.ascontiguousarray(first, dtype=np.int64)
    _, length, n = t.shape
    out = np.zeros((length, n), dtype=np.float32)
    lib().o64_path_dist_f32(_ptr(t), _ptr(r), length, n, _ptr(f), int(window), _ptr(out))
    return out


def trace64(charges, seeds, length, rule, scale=None, with_steps=False):
    """FP64 trajectory (PLAIN or ANGLE).  Returns rows[len, n, 3] (float64) [, steps]."""
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.float64)
    n = seeds.shape[0]
    rows = np.zeros((length, n, 3), dtype=np.float64)
    steps = np.zeros(n, dtype=np.uint32)
    sp = _params(rule)
    sc = rule.field_scale if scale is None else scale
    rc = lib().o64_trace(_ptr(charges), charges.shape[0], _ptr(seeds), n, length,
                         ctypes.byref(sp), sc, _ptr(rows), _ptr(steps))
    if rc:
        raise ValueError(f"o64_trace rc={rc}")
    return (rows, steps) if with_steps else rows


def trace64_jitter(charges, seeds, length, rule, rel, key, scale=None, with_steps=False):
    """trace64 with float32-sized rounding injected at every step: each field
    component and each new coordinate is scaled by 1 +- rel, signs drawn from
    ``key`` (deterministic).  Returns rows[len, n, 3] [, steps]."""
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.float64)
    n = seeds.shape[0]
    rows = np.zeros((length, n, 3), dtype=np.float64)
    steps = np.zeros(n, dtype=np.uint32)
    sp = _params(rule)
    sc = rule.field_scale if scale is None else scale
    rc = lib().o64_trace_jitter(_ptr(charges), charges.shape[0], _ptr(seeds), n, length,
                                ctypes.byref(sp), sc, _ptr(rows), _ptr(steps), float(rel), int(key))
    if rc:
        raise ValueError(f"o64_trace_jitter rc={rc}")
    return (rows, steps) if with_steps else rows


def verlet64(particles, charges, accel, box_min, box_max, dt, write_back=False):
    """FP64 particle step (Particle_System.cpp:108-190).  particles (n, 8)
    {x, y, z, q, vx, vy, vz, m}, charges (p, 4), accel (n, 3) = a(t).
    Returns (a(t + dt) (n, 3), new particles (n, 8) or None)."""
    parts = np.ascontiguousarray(particles, dtype=np.float64)
    ch = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1, 4)
    acc = np.array(accel, dtype=np.float64, order="C").reshape(-1, 3)
    bmin = np.ascontiguousarray(box_min, dtype=np.float64)
    bmax = np.ascontiguousarray(box_max, dtype=np.float64)
    out = np.zeros_like(parts) if write_back else None
    lib().o64_verlet_step(_ptr(parts), parts.shape[0], _ptr(ch) if ch.size else None, ch.shape[0],
                          _ptr(acc), _ptr(bmin), _ptr(bmax), float(dt),
                          _ptr(out) if out is not None else None)
    return acc, out


def adjoint64(charges, seeds, length, rule, G, scale=None):
    """FP64 gradient of sum(G * rows) w.r.t. charges (p,4) and seeds (n,3).

    ``G`` has the layout of ``trace64``'s rows: (len, n, 3)."""
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64)
    n = seeds.shape[0]
    gch = np.zeros_like(charges)
    gseed = np.zeros_like(seeds)
    sp = _params(rule)
    sc = rule.field_scale if scale is None else scale
    rc = lib().o64_adjoint(_ptr(charges), charges.shape[0], _ptr(seeds), n, length,
                           ctypes.byref(sp), sc, _ptr(G), _ptr(gch), _ptr(gseed))
    if rc:
        raise ValueError(f"o64_adjoint rc={rc}")
    return gch, gseed


def adjoint64_path(charges, seeds, length, rule, G, path_rows, path_steps, scale=None):
    """FP64 reverse sweep linearised along a given trajectory.

    ``path_rows``: (3, len, n) float32 (e.g. the GPU's rows); ``path_steps``:
    (n,) step counts.  G as in adjoint64.  Returns (grad_charges, grad_seeds)."""
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64)
    path = np.ascontiguousarray(np.asarray(path_rows, dtype=np.float32).transpose(1, 2, 0))
    st = np.ascontiguousarray(path_steps, dtype=np.uint32)
    n = seeds.shape[0]
    gch = np.zeros_like(charges)
    gseed = np.zeros_like(seeds)
    sp = _params(rule)
    sc = rule.field_scale if scale is None else scale
    rc = lib().o64_adjoint_path(_ptr(charges), charges.shape[0], _ptr(seeds), n, length,
                                ctypes.byref(sp), sc, _ptr(G), _ptr(gch), _ptr(gseed), _ptr(path),
                                _ptr(st))
    if rc:
        raise ValueError(f"o64_adjoint_path rc={rc}")
    return gch, gseed
