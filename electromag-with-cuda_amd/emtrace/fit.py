"""Inverse fit of charge configurations to traced field lines.

Given many configurations (a BatchLayout of charges and seeds) and target
lines (traced from the true charges), minimise

    L = sum_c mean_{lines, rows of c} |x - x_target|^2

over every charge's position and magnitude with Adam.  One iteration is a
batched forward with checkpoints, the loss gradient dL/dx, the batched adjoint
(dL/d charges) and an elementwise Adam update, all on one stream.  After one
eager warm-up iteration (which sizes every work buffer), the iteration is
captured into a single graph and replayed; the graph holds no memset and no
reduction (the loss value is computed outside it).  All kernels are
deterministic, so replays equal eager iterations bit for bit.
"""
from __future__ import annotations

import numpy as np

from . import api
from .rules import StepRule, angle


class InverseFit:
    def __init__(self, workloads, target_charges, rule: StepRule = angle(), ckpt_every: int = 50,
                 lr_pos: float = 5e-3, lr_q: float = 2.5e-3, betas=(0.9, 0.999), eps: float = 1e-12,
                 launch=None, order_lines: bool = True):
        import torch
        self.torch = torch
        self.rule = rule
        self.every = ckpt_every
        self.layout, charges, self.lines = api.pack_batch(workloads)
        if launch is None:
            # With line ends most lines are short and the few full-length lines
            # set the time of each pass: give every line as many lanes as the
            # largest configuration keeps busy (>= 32 charges per lane, <= 16),
            # in 128-thread work items (fit_64: 21.6-21.7 ms per iteration,
            # against 22.2-22.4 with 256 threads, 24.8 with 512, 25.6 with 64)
            S = 1
            while S < 16 and int(self.layout.p.max()) // (2 * S) >= 32:
                S *= 2
            launch = api.make_launch(S, 0, 128)
        self.theta = charges.clone()                       # parameters (sum p, 4)
        self.theta0 = charges.clone()
        self.true = torch.from_numpy(np.concatenate(target_charges)).cuda()
        self.plan = api.BatchPlan(self.layout, rule, launch, use_graph=False)
        # Lines that end early (the stop rule) idle their lanes while a wave
        # still holds a live line.  The loss is a sum over lines, so each
        # configuration's lines are put in order of their step count at the
        # starting charges (longest first, stable): waves then hold lines of
        # similar length, and the forward and adjoint skip waves whose lines have
        # all ended.  self.order[i][k] is the original index of line k of
        # configuration i.
        self.order = [np.arange(int(n)) for n in self.layout.n]
        if order_lines and rule.stop:
            self._order_lines()
        # targets: the same seeds traced through the true charges
        self.target = self.lines.clone()
        self.plan.run(self.true, self.target)
        self.ckpt = torch.empty(max(1, self.plan.ckpt_size(ckpt_every)), dtype=torch.float32,
                                device="cuda")
        self.grad = torch.zeros_like(self.theta)
        self.G = torch.zeros_like(self.lines)
        # per-element weight 2 / (n_c * len_c): d/dx of each configuration's mean
        wts = np.empty(self.layout.total_line_elems, dtype=np.float32)
        for i in range(len(self.layout.p)):
            o, n, ln = int(self.layout.line_off[i]), int(self.layout.n[i]), int(self.layout.len[i])
            wts[o:o + n * ln] = 2.0 / (n * ln)
        self.wts = torch.from_numpy(wts).cuda()
        # the same weight per charge: the adjoint is linear in dL/dx and a
        # configuration's charges see only its own lines, so the iteration
        # differentiates sum (x - x_target) and scales each charge's gradient
        # (one pass over the lines fewer than weighting dL/dx itself)
        cs = np.concatenate([np.full(int(self.layout.p[i]), 2.0 / (int(self.layout.n[i]) *
                                                                   int(self.layout.len[i])),
                                     dtype=np.float32) for i in range(len(self.layout.p))])
        self.cscale = torch.from_numpy(cs).cuda().unsqueeze(1)
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.t = torch.zeros((), dtype=torch.float32, device="cuda")
        self.lr = torch.tensor([lr_pos, lr_pos, lr_pos, lr_q], dtype=torch.float32, device="cuda")
        self.lr0 = self.lr.clone()
        self.b1, self.b2, self.eps = float(betas[0]), float(betas[1]), float(eps)
        self.graph = None

    def _order_lines(self):
        torch = self.torch
        steps = torch.empty(int(self.layout.n.sum()), dtype=torch.int32, device="cuda")
        probe = self.lines.clone()
        self.plan.run(self.theta, probe, steps)
        steps = steps.cpu().numpy()
        del probe
        o = 0
        for i in range(len(self.layout.n)):
            n = int(self.layout.n[i])
            perm = np.argsort(-steps[o:o + n].astype(np.int64), kind="stable")
            o += n
            view = self.layout.rows_view(self.lines, i)
            idx = torch.from_numpy(perm).cuda()
            view[:, 0, :] = view[:, 0, :].index_select(1, idx)
            self.order[i] = perm

    # one iteration, as a sequence of launches on the current stream
    def _body(self):
        torch = self.torch
        self.plan.run_ckpt(self.theta, self.lines, self.ckpt, self.every)
        torch.sub(self.lines, self.target, out=self.G)
        self.plan.adjoint(self.theta, self.lines, self.ckpt, self.every, self.G, self.grad)
        self.grad.mul_(self.cscale)
        self.t.add_(1.0)
        self.m.mul_(self.b1).add_(self.grad, alpha=1.0 - self.b1)
        self.v.mul_(self.b2).addcmul_(self.grad, self.grad, value=1.0 - self.b2)
        bc1 = 1.0 - torch.pow(torch.full_like(self.t, self.b1), self.t)
        bc2 = 1.0 - torch.pow(torch.full_like(self.t, self.b2), self.t)
        upd = (self.m / bc1) / ((self.v / bc2).sqrt() + self.eps)
        self.theta.sub_(upd * self.lr)

    def step_eager(self):
        self._body()

    def capture(self):
        """Warm up once (allocates the adjoint's work lists), then capture."""
        torch = self.torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._body()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()
        torch.cuda.synchronize()

    def step(self):
        if self.graph is None:
            self.capture()
        else:
            self.graph.replay()

    def set_lr_scale(self, scale: float):
        """Learning rates = the initial ones times ``scale`` from the next
        iteration on (an eager copy into the tensor the graph reads; e.g. a
        cosine decay over a fixed number of iterations)."""
        self.lr.copy_(self.lr0 * float(scale))

    # ---- reporting (outside the graph) ----
    def loss(self) -> float:
        torch = self.torch
        d = (self.lines - self.target) ** 2
        d = d.sum(dim=0) * (self.wts / 2.0)
        return float(d.sum().item())

    def recovery(self):
        """RMS position and magnitude errors vs the true charges, now and at
        start.  Field lines do not change when all charges of a configuration
        are scaled by one factor, so the magnitudes are observable only up to
        it: ``q_rms_scaled`` is the magnitude RMS after each configuration's
        magnitudes are scaled by the least-squares factor onto the truth."""
        torch = self.torch

        def rms(a, b):
            pos = torch.sqrt(((a[:, :3] - b[:, :3]) ** 2).sum(dim=1).mean()).item()
            mag = torch.sqrt(((a[:, 3] - b[:, 3]) ** 2).mean()).item()
            return pos, mag

        def scaled(a, b):
            sq, cnt = 0.0, 0
            for i in range(len(self.layout.p)):
                o, p = int(self.layout.charge_off[i]), int(self.layout.p[i])
                qa, qb = a[o:o + p, 3].double(), b[o:o + p, 3].double()
                f = (qa * qb).sum() / (qa * qa).sum()
                sq += float(((f * qa - qb) ** 2).sum().item())
                cnt += p
            return (sq / max(cnt, 1)) ** 0.5

        p0, q0 = rms(self.theta0, self.true)
        p1, q1 = rms(self.theta, self.true)
        return {"pos_rms_start": p0, "pos_rms": p1, "q_rms_start": q0, "q_rms": q1,
                "q_rms_scaled_start": scaled(self.theta0, self.true),
                "q_rms_scaled": scaled(self.theta, self.true)}
