"""torch.autograd binding of the tracer: forward = em_trace_dev_f32 with
checkpoints every ``ckpt_every`` steps, backward = em_adjoint_dev_f32 (segment
recompute from the checkpoints + reverse sweep + charge-gradient pass).

    rows = trace_lines(charges, seeds, length, rule, ckpt_every=50)
    loss = ((rows - target) ** 2).mean()
    loss.backward()          # charges.grad: (p, 4) = d/d{x, y, z, q}; seeds.grad: (n, 3)

``BatchTrace`` does the same for the packed configurations of a BatchPlan.
"""
from __future__ import annotations

import ctypes

import torch

from . import api
from ._lib import call
from .rules import StepRule, angle


class _TraceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, charges, seeds, length, rule, every, launch):
        charges = charges.detach().contiguous()
        n = seeds.shape[0]
        rows = api.seed_rows(seeds.detach().contiguous(), length)
        ckpt = torch.empty(max(1, api.ckpt_size(n, length, every)), dtype=torch.float32,
                           device=charges.device)
        api.trace_into(charges, rows, n, rule, launch, ckpt, every)
        ctx.save_for_backward(charges, rows, ckpt)
        ctx.rule, ctx.every, ctx.launch, ctx.n = rule, every, launch, n
        return rows

    @staticmethod
    def backward(ctx, grad_rows):
        charges, rows, ckpt = ctx.saved_tensors
        g = grad_rows.contiguous()
        _, length, pitch = rows.shape
        p = charges.shape[0]
        grad_c = torch.empty((p, 4), dtype=torch.float32, device=charges.device)
        grad_s = torch.empty((3, ctx.n), dtype=torch.float32, device=charges.device)
        st = api.make_step(ctx.rule)
        call("em_adjoint_dev_f32", charges.data_ptr(), p, rows[0].data_ptr(), rows[1].data_ptr(),
             rows[2].data_ptr(), ctx.n, length, pitch, ctypes.byref(st),
             api._launch_arg(ctx.launch), ckpt.data_ptr(), ctx.every, g[0].data_ptr(),
             g[1].data_ptr(), g[2].data_ptr(), grad_c.data_ptr(), grad_s.data_ptr(),
             api._stream_handle())
        return grad_c, grad_s.t(), None, None, None, None


def check_adjoint_buffers(charges, rows, ckpt, ckpt_every, grad_rows, n=None,
                          on_device: bool = True) -> None:
    """Validate a single-configuration adjoint call before any launch: charges
    (p, 4), rows and grad_rows (3, len, pitch) float32 and contiguous, grad_rows
    shaped like rows, ckpt_every >= 1 and a checkpoint buffer of at least
    em_ckpt_floats(n, len, ckpt_every) floats.  Raises TypeError / ValueError."""
    def tensor(t, name):
        if on_device:
            api._dev_check(t, name, torch.float32)
        elif not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous float32 tensor")

    for t, name in ((charges, "charges"), (rows, "rows"), (ckpt, "ckpt"), (grad_rows, "grad_rows")):
        tensor(t, name)
    if charges.dim() != 2 or charges.shape[1] != 4:
        raise ValueError("charges must be (p, 4)")
    if rows.dim() != 3 or rows.shape[0] != 3:
        raise ValueError("rows must be (3, len, pitch)")
    if tuple(grad_rows.shape) != tuple(rows.shape):
        raise ValueError(f"grad_rows must be shaped like rows {tuple(rows.shape)}")
    if ckpt_every < 1:
        raise ValueError("ckpt_every must be >= 1")
    _, length, pitch = rows.shape
    n = pitch if n is None else n
    if not 0 < n <= pitch:
        raise ValueError("need 0 < n <= pitch")
    need = api.ckpt_size(n, length, ckpt_every)
    if ckpt.numel() < need:
        raise ValueError(f"checkpoint buffer holds {ckpt.numel()} floats, needs {need}")


def trace_lines(charges: torch.Tensor, seeds: torch.Tensor, length: int,
                rule: StepRule = angle(), ckpt_every: int = 50, launch=None) -> torch.Tensor:
    """Differentiable trace: rows (3, len, n) with row 0 = seeds."""
    return _TraceFn.apply(charges, seeds, length, rule, ckpt_every, launch)


def adjoint(charges, rows, ckpt, ckpt_every, grad_rows, rule, launch=None, n=None):
    """Direct call of the adjoint (no autograd): returns (grad_charges, grad_seeds[3, n])."""
    check_adjoint_buffers(charges, rows, ckpt, ckpt_every, grad_rows, n)
    _, length, pitch = rows.shape
    n = pitch if n is None else n
    p = charges.shape[0]
    grad_c = torch.empty((p, 4), dtype=torch.float32, device=charges.device)
    grad_s = torch.empty((3, n), dtype=torch.float32, device=charges.device)
    st = api.make_step(rule)
    g = grad_rows.contiguous()
    call("em_adjoint_dev_f32", charges.data_ptr(), p, rows[0].data_ptr(), rows[1].data_ptr(),
         rows[2].data_ptr(), n, length, pitch, ctypes.byref(st), api._launch_arg(launch),
         ckpt.data_ptr(), ckpt_every, g[0].data_ptr(), g[1].data_ptr(), g[2].data_ptr(),
         grad_c.data_ptr(), grad_s.data_ptr(), api._stream_handle())
    return grad_c, grad_s


class _BatchFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, charges, plan, seed_lines, every):
        charges = charges.detach().contiguous()
        lines = seed_lines.detach().clone()
        ckpt = torch.empty(max(1, plan.ckpt_size(every)), dtype=torch.float32, device=charges.device)
        plan.run_ckpt(charges, lines, ckpt, every)
        ctx.save_for_backward(charges, lines, ckpt)
        ctx.plan, ctx.every = plan, every
        return lines

    @staticmethod
    def backward(ctx, grad_lines):
        charges, lines, ckpt = ctx.saved_tensors
        # charges beyond the layout's get no gradient from the adjoint: zero
        grad_c = torch.zeros_like(charges)
        ctx.plan.adjoint(charges, lines, ckpt, ctx.every, grad_lines.contiguous(), grad_c)
        return grad_c, None, None, None


def batch_trace(charges: torch.Tensor, plan: "api.BatchPlan", seed_lines: torch.Tensor,
                ckpt_every: int = 50) -> torch.Tensor:
    """Differentiable batched trace over a BatchPlan.  ``seed_lines`` is the
    packed (3, total) line buffer with each configuration's row 0 filled."""
    return _BatchFn.apply(charges, plan, seed_lines, ckpt_every)
