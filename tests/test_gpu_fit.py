"""Inverse fit (batched forward + adjoint + Adam in one captured graph).

* graph replays equal eager iterations bit for bit, with eager work and
  synchronisations between replays;
* the batched adjoint's gradients on a 4-configuration subset of the fit_64
  workload agree with the FP64 oracle: the FP64 sweep along the GPU's own path
  (all lines), and the FP64 adjoint of the FP64 trajectory on the lines that
  follow it;
* a short fit lowers the loss;
* the fit orders each configuration's lines by step count (longest first) and
  the order is a permutation of the seeds.
"""
import numpy as np
import pytest
import torch

from emtrace import gen, rules
from emtrace.fit import InverseFit
from helpers import fp64_sensitivity, rows64_to_soa

pytestmark = pytest.mark.gpu
# The FP64 adjoint of the FP64 trajectory is compared on lines that follow it
# (same step count, every point within FOLLOW_TOL) and are well conditioned:
# their FP64 path moves by at most SENS_TOL under float32-sized rounding
# (helpers.fp64_sensitivity).  On lines that amplify rounding more, the
# float32 and FP64 paths are linearised on measurably different trajectories.
# Measured on the 4-configuration subset: 0.9e-3 - 4.0e-3 with SENS_TOL = 1e-4
# (73-78 % of the lines), up to 8e-2 when lines up to 1e-3 are admitted.
FOLLOW_TOL = 1e-3
SENS_TOL = 1e-4


def _problem(count, p_range, n_range, length, sigma, seed):
    ws = gen.batched_configs(count, p_range, n_range, length, seed)
    truth = [gen.perturbed(w.charges, sigma, seed=seed * 100 + i) for i, w in enumerate(ws)]
    return ws, truth


def test_graph_replay_equals_eager_iterations(gpu):
    ws, truth = _problem(5, (64, 400), (64, 300), 60, 0.1, 3)
    eager = InverseFit(ws, truth, rules.angle(), ckpt_every=9)
    want = []
    for _ in range(4):
        eager.step_eager()
        torch.cuda.synchronize()
        want.append(eager.theta.clone())
    graphed = InverseFit(ws, truth, rules.angle(), ckpt_every=9)
    graphed.step()  # warm-up iteration (eager) + capture
    torch.cuda.synchronize()
    assert torch.equal(graphed.theta, want[0])
    for k in range(1, 4):
        graphed.step()  # replay
        torch.cuda.synchronize()
        _ = graphed.loss()  # eager work between replays
        assert torch.equal(graphed.theta, want[k]), k


def test_fit_64_subset_gradients_vs_fp64(gpu, orc):
    """First 4 configurations of the fit_64 workload (800 rows, sigma 0.1)."""
    ws_all = gen.fit_64()
    ws = ws_all[:4]
    truth = [gen.perturbed(w.charges, 0.1, seed=i) for i, w in enumerate(ws)]
    fit = InverseFit(ws, truth, rules.angle(), ckpt_every=50)
    fit.plan.run_ckpt(fit.theta, fit.lines, fit.ckpt, fit.every)
    torch.sub(fit.lines, fit.target, out=fit.G)
    fit.G.mul_(fit.wts)
    fit.plan.adjoint(fit.theta, fit.lines, fit.ckpt, fit.every, fit.G, fit.grad)
    steps = torch.empty(int(fit.layout.n.sum()), dtype=torch.int32, device="cuda")
    fit.plan.run(fit.theta, fit.lines, steps)
    torch.cuda.synchronize()
    grad = fit.grad.cpu().numpy().astype(np.float64)
    lines = fit.lines
    G = fit.G
    steps = steps.cpu().numpy().astype(np.int64)
    rep = []
    off_line = 0
    for i, w in enumerate(ws):
        rows = fit.layout.rows_view(lines, i).cpu().numpy()
        g_rows = fit.layout.rows_view(G, i).cpu().numpy().astype(np.float64)
        st = steps[off_line:off_line + w.n]
        off_line += w.n
        G64 = np.ascontiguousarray(g_rows.transpose(1, 2, 0))
        seeds = np.ascontiguousarray(rows[:, 0, :].T)  # the fit's own line order
        assert np.array_equal(np.sort(seeds, axis=0), np.sort(w.seeds, axis=0))
        gp, _ = orc.adjoint64_path(w.charges, seeds, w.len, rules.angle(), G64, rows, st)
        o = int(fit.layout.charge_off[i])
        gg = grad[o:o + w.p]
        r_path = np.linalg.norm(gg - gp) / np.linalg.norm(gp)
        # own FP64 trajectory, on the lines that follow it
        r64, s64 = orc.trace64(w.charges, seeds, w.len, rules.angle(), with_steps=True)
        err = np.abs(rows.astype(np.float64) - rows64_to_soa(r64)).max(axis=(0, 1))
        follow = (st == s64.astype(np.int64)) & (err <= FOLLOW_TOL)
        first = np.full(w.n, w.len)
        sens = np.zeros(w.n)
        idx = np.nonzero(follow)[0]
        sens[idx] = fp64_sensitivity(w.charges, seeds, w.len, rules.angle(), first, idx)
        ok = follow & (sens <= SENS_TOL)
        Gm = G64.copy()
        Gm[:, ~ok, :] = 0.0
        g64, _ = orc.adjoint64(w.charges, seeds, w.len, rules.angle(), Gm)
        Gmf = np.ascontiguousarray(Gm.transpose(2, 0, 1)).astype(np.float32)
        rep.append((w.p, w.n, float(r_path), float(ok.mean()), g64, Gmf, i))
        print(f"[fit_64 subset] cfg {i}: p={w.p} n={w.n} rel(gpu, FP64 along gpu path) {r_path:.2e}; "
              f"lines following FP64 {follow.mean():.3f}, and well conditioned {ok.mean():.3f}",
              flush=True)
        assert r_path < 1e-3
        assert follow.mean() > 0.8 and ok.mean() > 0.6
    # the masked gradients (agreeing lines only) through the batched GPU adjoint
    for (p, n, _, _, g64, Gmf, i) in rep:
        fit.layout.rows_view(fit.G, i).copy_(torch.from_numpy(Gmf).cuda())
    fit.plan.adjoint(fit.theta, fit.lines, fit.ckpt, fit.every, fit.G, fit.grad)
    torch.cuda.synchronize()
    grad = fit.grad.cpu().numpy().astype(np.float64)
    for (p, n, _, _, g64, _, i) in rep:
        o = int(fit.layout.charge_off[i])
        r = np.linalg.norm(grad[o:o + p] - g64) / np.linalg.norm(g64)
        print(f"[fit_64 subset] cfg {i}: rel(gpu, FP64 adjoint of FP64 trajectory) {r:.2e}", flush=True)
        assert r < 1e-2


def test_scaled_gradient_equals_weighted_dldx(gpu):
    """The iteration differentiates sum (x - x_target) and scales each charge's
    gradient by its configuration's weight 2 / (n len); the adjoint of the
    weighted dL/dx gives the same gradient up to float rounding."""
    ws = gen.fit_64()[:6]
    truth = [gen.perturbed(w.charges, 0.1, seed=i) for i, w in enumerate(ws)]
    fit = InverseFit(ws, truth, rules.angle(), ckpt_every=50)
    fit.plan.run_ckpt(fit.theta, fit.lines, fit.ckpt, fit.every)
    torch.sub(fit.lines, fit.target, out=fit.G)
    fit.plan.adjoint(fit.theta, fit.lines, fit.ckpt, fit.every, fit.G, fit.grad)
    scaled = (fit.grad * fit.cscale).cpu().numpy().astype(np.float64)
    fit.G.mul_(fit.wts)
    fit.plan.adjoint(fit.theta, fit.lines, fit.ckpt, fit.every, fit.G, fit.grad)
    torch.cuda.synchronize()
    weighted = fit.grad.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(scaled - weighted) / np.linalg.norm(weighted)
    print(f"[fit] scaled vs weighted gradient: rel {rel:.2e}", flush=True)
    assert rel < 1e-5


def test_short_fit_lowers_the_loss(gpu):
    ws, truth = _problem(8, (128, 512), (256, 512), 200, 0.1, 5)
    fit = InverseFit(ws, truth, rules.angle(), ckpt_every=50)
    fit.plan.run(fit.theta, fit.lines)
    l0 = fit.loss()
    for _ in range(30):
        fit.step()
    fit.plan.run(fit.theta, fit.lines)
    l1 = fit.loss()
    rec = fit.recovery()
    print(f"[fit] loss {l0:.4g} -> {l1:.4g}; recovery {rec}", flush=True)
    assert l1 < l0


def test_lines_ordered_by_step_count(gpu):
    ws, truth = _problem(6, (64, 400), (64, 300), 120, 0.1, 8)
    fit = InverseFit(ws, truth, rules.angle(), ckpt_every=20)
    steps = torch.empty(int(fit.layout.n.sum()), dtype=torch.int32, device="cuda")
    fit.plan.run(fit.theta, fit.lines, steps)
    steps = steps.cpu().numpy()
    o = 0
    for i, w in enumerate(ws):
        st = steps[o:o + w.n]
        o += w.n
        assert (np.diff(st.astype(np.int64)) <= 0).all(), i  # longest first
        seeds = fit.layout.rows_view(fit.lines, i)[:, 0, :].t().cpu().numpy()
        assert np.array_equal(seeds, w.seeds[fit.order[i]]), i


def test_well_posed_fit_recovers_the_charges(gpu):
    """A well-posed inverse problem -- 64 configurations of 2 charges, 256
    lines of 60 rows, positions perturbed by N(0, 0.02^2) (a fifth of the seed
    radius), the plain rule (row s sits at arc length s / resolution) -- is
    recovered by the same machinery: 200 captured iterations with a cosine-
    decayed learning rate bring the position RMS error down by more than 2x
    and the magnitude error (modulo each configuration's overall scale, which
    the lines cannot see) by more than 1.5x."""
    import math
    ws = gen.batched_configs(64, (2, 2), (256, 256), 60, 5)
    truth = [gen.perturbed(w.charges, 0.02, seed=100 + i) for i, w in enumerate(ws)]
    fit = InverseFit(ws, truth, rules.plain(10.0), ckpt_every=20, lr_pos=4e-3, lr_q=2e-3)
    for it in range(200):
        fit.set_lr_scale(0.5 * (1.0 + math.cos(math.pi * it / 200)))
        fit.step()
    torch.cuda.synchronize()
    r = fit.recovery()
    print(f"[fit well-posed] {r}", flush=True)
    assert r["pos_rms"] < 0.5 * r["pos_rms_start"]
    assert r["q_rms_scaled"] < r["q_rms_scaled_start"] / 1.5
