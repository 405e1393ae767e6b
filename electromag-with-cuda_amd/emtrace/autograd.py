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


def trace_lines(charges: torch.Tensor, seeds: torch.Tensor, length: int,
                rule: StepRule = angle(), ckpt_every: int = 50, launch=None) -> torch.Tensor:
    """Differentiable trace: rows (3, len, n) with row 0 = seeds."""
    return _TraceFn.apply(charges, seeds, length, rule, ckpt_every, launch)


def adjoint(charges, rows, ckpt, ckpt_every, grad_rows, rule, launch=None, n=None):
    """Direct call of the adjoint (no autograd): returns (grad_charges, grad_seeds[3, n])."""
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
        grad_c = torch.empty_like(charges)
        ctx.plan.adjoint(charges, lines, ckpt, ctx.every, grad_lines.contiguous(), grad_c)
        return grad_c, None, None, None


def batch_trace(charges: torch.Tensor, plan: "api.BatchPlan", seed_lines: torch.Tensor,
                ckpt_every: int = 50) -> torch.Tensor:
    """Differentiable batched trace over a BatchPlan.  ``seed_lines`` is the
    packed (3, total) line buffer with each configuration's row 0 filled."""
    return _BatchFn.apply(charges, plan, seed_lines, ckpt_every)
