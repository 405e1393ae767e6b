"""Python-level API over the C ABI (numpy for host arrays, torch for device memory).

Host-array calls mirror the reference's calc-lines interface
(``CalcField_CPU<float>``, ElectroMag/src/CPU Implement.h:25-30): charges are
``(p, 4)`` float32 records, lines are three step-major ``len*n`` arrays with the
seeds in row 0.  Here they are carried as one ``(3, len, n)`` array.

Device calls take torch tensors on the current HIP device and run on torch's
current stream; torch is only the allocator and stream provider.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import call, check, em_batch_cfg, em_launch, em_perf, em_step, lib
from .rules import ANGLE, PLAIN, REF, StepRule


def make_step(rule: StepRule) -> em_step:
    return em_step(int(rule.rule), float(rule.resolution), float(rule.h0), float(rule.hmin),
                   float(rule.hmax), float(rule.turn), int(bool(rule.stop)))


def make_launch(lanes_per_line: int = 0, tile: int = 0, block: int = 0,
                unroll: int = 0) -> em_launch:
    """A launch shape; 0 leaves a field to the library.  ``unroll`` (2, 4 or 8
    charge records per pipelined group) changes the speed only, never the bits."""
    return em_launch(int(lanes_per_line), int(tile), int(block), int(unroll))


# Line ends (angle rule with stop) without checkpoints over ENDS_MIN_CHARGES to
# WIDE_MAX_CHARGES charges take the register-resident trace, which has one
# fixed launch: a launch shape given for it is not used.
ENDS_MIN_CHARGES = 4096
WIDE_MAX_CHARGES = 16384


def takes_launch_shape(p: int, rule, ckpt: bool = False) -> bool:
    """Whether a single-configuration trace of p charges with ``rule`` uses a
    launch shape (all but the register-resident line-ends trace)."""
    return not (rule.stop and rule.rule == 2 and not ckpt
                and ENDS_MIN_CHARGES <= p <= WIDE_MAX_CHARGES)


def _launch_arg(launch):
    if launch is None:
        return None
    if isinstance(launch, em_launch):
        return ctypes.byref(launch)
    if isinstance(launch, dict):
        return ctypes.byref(make_launch(**launch))
    return ctypes.byref(make_launch(*launch))


def _np_ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ---------------------------------------------------------------------------
# host arrays
def calc_lines(charges: np.ndarray, x: np.ndarray, y: np.ndarray, z: np.ndarray, n: int,
               resolution: float, use_curvature: bool = False):
    """Drop-in for ``CalcField_CPU<float>``: fills rows 1.. of x/y/z in place.

    Returns ``(status, perf)`` -- the status is the reference's int code
    (0 ok, 1 n == 0, 2 resolution == 0, 3 fewer than two rows)."""
    charges = np.ascontiguousarray(charges, dtype=np.float32)
    for a in (x, y, z):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise TypeError("x, y, z must be contiguous float32 arrays (written in place)")
    length = x.size // n if n else 0
    perf = em_perf()
    rc = lib().em_calc_lines_f32(_np_ptr(charges), charges.shape[0], _np_ptr(x), _np_ptr(y),
                                 _np_ptr(z), n, length, float(resolution), int(bool(use_curvature)),
                                 ctypes.byref(perf))
    return rc, perf.as_dict()


def trace_host(charges: np.ndarray, seeds: np.ndarray, length: int, rule: StepRule, launch=None):
    """Trace on the GPU from host arrays.  Returns (rows[3, len, n], perf dict, steps[n])."""
    charges = np.ascontiguousarray(charges, dtype=np.float32)
    seeds = np.asarray(seeds, dtype=np.float32)
    n = seeds.shape[0]
    rows = np.zeros((3, length, n), dtype=np.float32)
    rows[:, 0, :] = seeds.T
    steps = np.zeros(n, dtype=np.uint32)
    perf = em_perf()
    st = make_step(rule)
    call("em_trace_host_f32", _np_ptr(charges), charges.shape[0], _np_ptr(rows[0]),
         _np_ptr(rows[1]), _np_ptr(rows[2]), n, length, ctypes.byref(st), _launch_arg(launch),
         ctypes.byref(perf), _np_ptr(steps))
    return rows, perf.as_dict(), steps


def calc_lines64(charges: np.ndarray, x: np.ndarray, y: np.ndarray, z: np.ndarray, n: int,
                 resolution: float, use_curvature: bool = False):
    """Drop-in for ``CalcField_CPU<double>`` (IEEE binary64 kernel); as calc_lines."""
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    for a in (x, y, z):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError("x, y, z must be contiguous float64 arrays (written in place)")
    length = x.size // n if n else 0
    perf = em_perf()
    rc = lib().em_calc_lines_f64(_np_ptr(charges), charges.shape[0], _np_ptr(x), _np_ptr(y),
                                 _np_ptr(z), n, length, float(resolution), int(bool(use_curvature)),
                                 ctypes.byref(perf))
    return rc, perf.as_dict()


def trace_host64(charges: np.ndarray, seeds: np.ndarray, length: int, rule: StepRule):
    """FP64 trace from host arrays.  Returns (rows[3, len, n] float64, perf, steps[n])."""
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    seeds = np.asarray(seeds, dtype=np.float64)
    n = seeds.shape[0]
    rows = np.zeros((3, length, n), dtype=np.float64)
    rows[:, 0, :] = seeds.T
    steps = np.zeros(n, dtype=np.uint32)
    perf = em_perf()
    st = make_step(rule)
    call("em_trace_host_f64", _np_ptr(charges), charges.shape[0], _np_ptr(rows[0]),
         _np_ptr(rows[1]), _np_ptr(rows[2]), n, length, ctypes.byref(st), ctypes.byref(perf),
         _np_ptr(steps))
    return rows, perf.as_dict(), steps


def verlet_step(particles: np.ndarray, charges: np.ndarray, accel: np.ndarray, box_min, box_max,
                dt: float, write_back: bool = False):
    """The particle step (emgVerletStep, Particle_System.cpp:108-190) on the GPU
    from host arrays.  particles (n, 8) float32 {x, y, z, q, vx, vy, vz, m};
    charges (p, 4); accel (n, 3) = a(t), overwritten with a(t + dt); with
    write_back the particles are advanced in place.  Returns perf."""
    for a, name, w in ((particles, "particles", 8), (accel, "accel", 3)):
        if a.dtype != np.float32 or not a.flags.c_contiguous or a.ndim != 2 or a.shape[1] != w:
            raise TypeError(f"{name} must be a contiguous float32 (n, {w}) array (written in place)")
    if accel.shape[0] != particles.shape[0]:
        raise ValueError("accel must have one row per particle")
    charges = np.ascontiguousarray(charges, dtype=np.float32).reshape(-1, 4)
    bmin = (ctypes.c_float * 3)(*[float(v) for v in box_min])
    bmax = (ctypes.c_float * 3)(*[float(v) for v in box_max])
    perf = em_perf()
    call("em_verlet_step_f32", _np_ptr(particles), particles.shape[0],
         _np_ptr(charges) if charges.size else None, charges.shape[0], _np_ptr(accel), bmin, bmax,
         float(dt), int(bool(write_back)), ctypes.byref(perf))
    return perf.as_dict()


# ---------------------------------------------------------------------------
# device tensors
def _torch():
    import torch
    return torch


def _stream_handle():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_check(t, name, dtype=None):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}")


def seed_rows(seeds, length: int, pitch: Optional[int] = None):
    """Allocate a (3, len, pitch) float32 device tensor with the seeds in row 0."""
    torch = _torch()
    n = seeds.shape[0]
    pitch = n if pitch is None else pitch
    rows = torch.empty((3, length, pitch), dtype=torch.float32, device=seeds.device)
    rows[:, 0, :n] = seeds.t()
    return rows


def ckpt_size(n: int, length: int, every: int) -> int:
    return int(lib().em_ckpt_floats(n, length, every))


def trace_into(charges, rows, n: int, rule: StepRule, launch=None, ckpt=None, ckpt_every: int = 0,
               steps=None, noalloc: bool = False):
    """Trace in place on device rows (3, len, pitch); row 0 must hold the seeds.
    ``steps`` (optional int32 device tensor of n) receives each line's step count.
    ``noalloc`` calls ``em_trace_dev_f32_noalloc``: the same rows, with the
    line-ends work memory taken from the library's fixed per-device array
    (no device allocation; larger problems raise EmError)."""
    torch = _torch()
    _dev_check(charges, "charges", torch.float32)
    _dev_check(rows, "rows", torch.float32)
    if charges.dim() != 2 or charges.shape[1] != 4:
        raise ValueError("charges must be (p, 4)")
    if rows.dim() != 3 or rows.shape[0] != 3:
        raise ValueError("rows must be (3, len, pitch)")
    _, length, pitch = rows.shape
    st = make_step(rule)
    cp = None
    if ckpt is not None:
        _dev_check(ckpt, "ckpt", torch.float32)
        if ckpt.numel() < ckpt_size(n, length, ckpt_every):
            raise ValueError("checkpoint buffer too small")
        cp = ckpt.data_ptr()
    sp = None
    if steps is not None:
        _dev_check(steps, "steps", torch.int32)
        if steps.numel() < n:
            raise ValueError("steps buffer too small")
        sp = steps.data_ptr()
    if not 0 < n <= pitch:
        raise ValueError("need 0 < n <= pitch")
    call("em_trace_dev_f32_noalloc" if noalloc else "em_trace_dev_f32", charges.data_ptr(),
         charges.shape[0], rows[0].data_ptr(), rows[1].data_ptr(), rows[2].data_ptr(), n, length,
         pitch, ctypes.byref(st), _launch_arg(launch), cp, ckpt_every, sp, _stream_handle())
    return rows


def trace(charges, seeds, length: int, rule: StepRule, launch=None, with_steps: bool = False):
    """Trace from device seeds (n, 3).  Returns rows (3, len, n) on the device
    [and the int32 step count of every line]."""
    rows = seed_rows(seeds, length)
    if not with_steps:
        return trace_into(charges, rows, seeds.shape[0], rule, launch)
    torch = _torch()
    steps = torch.empty(seeds.shape[0], dtype=torch.int32, device=seeds.device)
    trace_into(charges, rows, seeds.shape[0], rule, launch, steps=steps)
    return rows, steps


def trace64(charges, seeds, length: int, rule: StepRule, with_steps: bool = False):
    """FP64 trace from device seeds (n, 3) float64; charges (p, 4) float64.
    Returns float64 rows (3, len, n) [and int32 step counts]."""
    torch = _torch()
    _dev_check(charges, "charges", torch.float64)
    if charges.dim() != 2 or charges.shape[1] != 4:
        raise ValueError("charges must be (p, 4)")
    n = seeds.shape[0]
    rows = torch.empty((3, length, n), dtype=torch.float64, device=seeds.device)
    rows[:, 0, :] = seeds.t()
    steps = torch.empty(n, dtype=torch.int32, device=seeds.device) if with_steps else None
    st = make_step(rule)
    call("em_trace_dev_f64", charges.data_ptr(), charges.shape[0], rows[0].data_ptr(),
         rows[1].data_ptr(), rows[2].data_ptr(), n, length, n, ctypes.byref(st),
         steps.data_ptr() if steps is not None else None, _stream_handle())
    return (rows, steps) if with_steps else rows


def verlet_step_dev(particles, charges, accel, box_min, box_max, dt: float, out=None):
    """Device form of verlet_step: particles (n, 8), charges (p, 4), accel (n, 3)
    float32 tensors; accel is updated in place; ``out`` (n, 8), if given, receives
    the advanced particles (it must not be ``particles``)."""
    torch = _torch()
    _dev_check(particles, "particles", torch.float32)
    _dev_check(accel, "accel", torch.float32)
    if charges.numel():
        _dev_check(charges, "charges", torch.float32)
    if out is not None:
        _dev_check(out, "out", torch.float32)
    bmin = (ctypes.c_float * 3)(*[float(v) for v in box_min])
    bmax = (ctypes.c_float * 3)(*[float(v) for v in box_max])
    call("em_verlet_step_dev_f32", particles.data_ptr(), particles.shape[0],
         charges.data_ptr() if charges.numel() else None, charges.shape[0] if charges.numel() else 0,
         accel.data_ptr(), bmin, bmax, float(dt), out.data_ptr() if out is not None else None,
         _stream_handle())
    return accel


def field(charges, points, scale: float = 1.0, launch=None):
    """E at device points (m, 3) -> (m, 3): scale * sum q (x - c)/|x - c|^3."""
    torch = _torch()
    _dev_check(charges, "charges", torch.float32)
    pts = points.t().contiguous().float()
    m = pts.shape[1]
    E = torch.empty((3, m), dtype=torch.float32, device=points.device)
    call("em_field_dev_f32", charges.data_ptr(), charges.shape[0], pts[0].data_ptr(),
         pts[1].data_ptr(), pts[2].data_ptr(), m, float(scale), E[0].data_ptr(),
         E[1].data_ptr(), E[2].data_ptr(), _launch_arg(launch), _stream_handle())
    return E.t()


# ---------------------------------------------------------------------------
# batched form
@dataclass
class BatchLayout:
    """Packing of many configurations: charges (sum p, 4); lines (3, total)."""
    p: np.ndarray
    n: np.ndarray
    len: np.ndarray
    charge_off: np.ndarray
    line_off: np.ndarray
    total_charges: int
    total_line_elems: int

    @staticmethod
    def of(ps: Sequence[int], ns: Sequence[int], lens: Sequence[int]) -> "BatchLayout":
        p = np.asarray(ps, dtype=np.int64)
        n = np.asarray(ns, dtype=np.int64)
        ln = np.asarray(lens, dtype=np.int64)
        coff = np.concatenate([[0], np.cumsum(p)[:-1]])
        sizes = n * ln
        loff = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        return BatchLayout(p, n, ln, coff, loff, int(p.sum()), int(sizes.sum()))

    def cfg_array(self):
        arr = (em_batch_cfg * len(self.p))()
        for i in range(len(self.p)):
            arr[i] = em_batch_cfg(int(self.charge_off[i]), int(self.line_off[i]), int(self.n[i]),
                                  int(self.p[i]), int(self.n[i]), int(self.len[i]), 0)
        return arr

    def rows_view(self, lines, i: int):
        """(3, len_i, n_i) view of configuration i inside the packed lines."""
        o, n, ln = int(self.line_off[i]), int(self.n[i]), int(self.len[i])
        return lines[:, o:o + n * ln].reshape(3, ln, n)

    def interactions(self) -> int:
        return int((self.p * self.n * (self.len - 1)).sum())


def check_batch_buffers(layout: BatchLayout, charges, lines, *, steps=None, ckpt=None,
                        ckpt_floats: Optional[int] = None, grad_lines=None, grad_charges=None,
                        on_device: bool = True) -> None:
    """Validate the packed buffers of a batch call before any launch: float32
    (int32 for steps), contiguous, and at least as large as the layout needs
    -- charges (>= sum p, 4), lines and grad_lines (3, >= total line elements),
    steps (>= sum n), ckpt (>= ckpt_floats), grad_charges (>= sum p, 4).
    Raises TypeError / ValueError.  ``on_device`` also requires GPU tensors."""
    torch = _torch()

    def tensor(t, name, dtype):
        if on_device:
            _dev_check(t, name, dtype)
        else:
            if not isinstance(t, torch.Tensor):
                raise TypeError(f"{name} must be a torch tensor")
            if not t.is_contiguous():
                raise ValueError(f"{name} must be contiguous")
            if t.dtype != dtype:
                raise TypeError(f"{name} must be {dtype}")

    def records(t, name):
        tensor(t, name, torch.float32)
        if t.dim() != 2 or t.shape[1] != 4 or t.shape[0] < layout.total_charges:
            raise ValueError(f"{name} must be (>= {layout.total_charges}, 4), got {tuple(t.shape)}")

    def planes(t, name):
        tensor(t, name, torch.float32)
        if t.dim() != 2 or t.shape[0] != 3 or t.shape[1] < layout.total_line_elems:
            raise ValueError(f"{name} must be (3, >= {layout.total_line_elems}), got {tuple(t.shape)}")

    records(charges, "charges")
    planes(lines, "lines")
    if steps is not None:
        tensor(steps, "steps", torch.int32)
        if steps.numel() < int(layout.n.sum()):
            raise ValueError("steps buffer smaller than the layout's lines")
    if ckpt is not None:
        tensor(ckpt, "ckpt", torch.float32)
        if ckpt_floats is None or ckpt.numel() < ckpt_floats:
            raise ValueError(f"checkpoint buffer holds {ckpt.numel()} floats, needs {ckpt_floats}")
    if grad_lines is not None:
        planes(grad_lines, "grad_lines")
    if grad_charges is not None:
        records(grad_charges, "grad_charges")


class BatchPlan:
    """Persistent batched launch over a BatchLayout (optionally hipGraph-replayed)."""

    def __init__(self, layout: BatchLayout, rule: StepRule, launch=None, use_graph: bool = True):
        self.layout = layout
        self._cfgs = layout.cfg_array()
        self._step = make_step(rule)
        h = ctypes.c_void_p()
        call("em_batch_create", self._cfgs, len(layout.p), ctypes.byref(self._step),
             _launch_arg(launch), int(use_graph), ctypes.byref(h))
        self._h = h

    def info(self):
        inter = ctypes.c_double()
        items = ctypes.c_int32()
        grid = ctypes.c_int32()
        used = em_launch()
        call("em_batch_info", self._h, ctypes.byref(inter), ctypes.byref(items), ctypes.byref(grid),
             ctypes.byref(used))
        return {"interactions": inter.value, "items": items.value, "grid": grid.value,
                "lanes_per_line": used.lanes_per_line, "tile": used.tile, "block": used.block,
                "unroll": used.unroll}

    def run(self, charges, lines, steps=None):
        """Trace every configuration; ``steps`` (optional int32 device tensor of
        sum(n)) receives the step counts, configurations back to back."""
        check_batch_buffers(self.layout, charges, lines, steps=steps)
        call("em_batch_run", self._h, charges.data_ptr(), lines[0].data_ptr(), lines[1].data_ptr(),
             lines[2].data_ptr(), None if steps is None else steps.data_ptr(), _stream_handle())

    def ckpt_size(self, every: int) -> int:
        return int(lib().em_batch_ckpt_floats(self._h, every))

    def run_ckpt(self, charges, lines, ckpt, every: int, steps=None):
        """run() that also stores checkpoints every ``every`` steps into ``ckpt``
        (at least ckpt_size(every) floats)."""
        if every < 1:
            raise ValueError("every must be >= 1")
        check_batch_buffers(self.layout, charges, lines, steps=steps, ckpt=ckpt,
                            ckpt_floats=self.ckpt_size(every))
        call("em_batch_run_ckpt", self._h, charges.data_ptr(), lines[0].data_ptr(),
             lines[1].data_ptr(), lines[2].data_ptr(), ckpt.data_ptr(), every,
             None if steps is None else steps.data_ptr(), _stream_handle())

    def adjoint(self, charges, lines, ckpt, every: int, grad_lines, grad_charges):
        """dL/d(charges) into grad_charges (>= sum p, 4) from dL/d(rows) in
        grad_lines (the packed line layout) and run_ckpt's checkpoints."""
        if every < 1:
            raise ValueError("every must be >= 1")
        check_batch_buffers(self.layout, charges, lines, ckpt=ckpt,
                            ckpt_floats=self.ckpt_size(every), grad_lines=grad_lines,
                            grad_charges=grad_charges)
        call("em_batch_adjoint", self._h, charges.data_ptr(), lines[0].data_ptr(),
             lines[1].data_ptr(), lines[2].data_ptr(), ckpt.data_ptr(), every,
             grad_lines[0].data_ptr(), grad_lines[1].data_ptr(), grad_lines[2].data_ptr(),
             grad_charges.data_ptr(), _stream_handle())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().em_batch_destroy(self._h), "em_batch_destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_batch(workloads, device=None):
    """Pack a list of gen.Workload into (layout, charges, lines) device tensors."""
    torch = _torch()
    layout = BatchLayout.of([w.p for w in workloads], [w.n for w in workloads],
                            [w.len for w in workloads])
    charges = torch.from_numpy(np.concatenate([w.charges for w in workloads])).to(device or "cuda")
    lines = torch.empty((3, layout.total_line_elems), dtype=torch.float32, device=charges.device)
    for i, w in enumerate(workloads):
        v = layout.rows_view(lines, i)
        v[:, 0, :] = torch.from_numpy(w.seeds.T.copy()).to(charges.device)
    return layout, charges, lines
