"""GPU adjoint vs the FP64 CPU adjoint and FP64 central finite differences.

Loss (differentiation workload): L = mean over points of |x - x_target|^2, the
targets traced from the same seeds with charges perturbed (positions +
N(0, sigma^2), magnitudes * U[0.95, 1.05]).  The gradient is checked against
  * the FP64 adjoint (oracle/oracle64.c: o64_adjoint) with the same dL/dx,
  * FP64 central differences (h = 1e-4) of the FP64 trace, on lines whose
    step counts stay fixed under every perturbation (the stop rule makes the
    loss jump where a line ends one step earlier or later) and on sampled
    parameters whose difference quotient has converged (FD(1e-4) within 1 % of
    FD(1e-5): no clamp kink inside +-h) -- at least 10 of 24,
with the tolerance written in each test.  The comparisons use the lines whose
float32 trace follows the FP64 trace (same step count, points within 1e-3):
elsewhere the two took different discrete decisions (where a line stops, which
clamp applies) and their gradients belong to different smooth pieces.
"""
import numpy as np
import pytest
import torch

from emtrace import gen, rules
from helpers import rows64_to_soa

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def forward_and_adjoint(em, charges, seeds, length, rule, G, every):
    """GPU forward (with checkpoints) and adjoint of sum(G . rows)."""
    c = dev(charges)
    n = seeds.shape[0]
    rows = em.seed_rows(dev(seeds), length)
    ck = torch.empty(max(1, em.ckpt_size(n, length, every)), dtype=torch.float32, device="cuda")
    em.trace_into(c, rows, n, rule, None, ck, every)
    gc, gs = em.adjoint(c, rows, ck, every, dev(G.astype(np.float32)), rule)
    torch.cuda.synchronize()
    return rows.cpu().numpy(), gc.cpu().numpy().astype(np.float64), gs.cpu().numpy().T.astype(np.float64)


def forward_steps(em, charges, seeds, length, rule, every=50):
    """Step counts of the checkpointed forward (the one-kernel path that the
    adjoint's recompute repeats; without checkpoints, line ends of many charges
    are traced by kernels whose charge sums are split differently)."""
    n = seeds.shape[0]
    rows = em.seed_rows(dev(seeds), length)
    ck = torch.empty(max(1, em.ckpt_size(n, length, every)), dtype=torch.float32, device="cuda")
    steps = torch.empty(n, dtype=torch.int32, device="cuda")
    em.trace_into(dev(charges), rows, n, rule, None, ck, every, steps=steps)
    torch.cuda.synchronize()
    return steps.cpu().numpy().astype(np.int64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("rule", [rules.plain(10.0), rules.angle()], ids=["plain", "angle"])
def test_adjoint_matches_fp64_adjoint_small(gpu, orc, rule):
    """(1) vs the FP64 sweep linearised along the GPU's own path (same discrete
    decisions): the float arithmetic of the adjoint; (2) vs the FP64 adjoint of
    the FP64 trajectory, on lines that follow it."""
    w = gen.plain_config(40, 96, 4.0, 60, seed=21)
    rng = np.random.default_rng(0)
    G = rng.normal(size=(3, w.len, w.n))
    g_steps = forward_steps(gpu, w.charges, w.seeds, w.len, rule, 7)
    rows, gc, gs = forward_and_adjoint(gpu, w.charges, w.seeds, w.len, rule, G, 7)
    G64 = np.ascontiguousarray(G.transpose(1, 2, 0))
    gp, gsp = orc.adjoint64_path(w.charges, w.seeds, w.len, rule, G64, rows, g_steps)
    r64, s64 = orc.trace64(w.charges, w.seeds, w.len, rule, with_steps=True)
    ok = agreeing_lines(rows, g_steps, rows64_to_soa(r64), s64.astype(np.int64))
    gc64, gs64 = orc.adjoint64(w.charges, w.seeds, w.len, rule, G64)
    print(f"[adjoint small {rules.NAMES[rule.rule]}] rel vs FP64 along GPU path: charges "
          f"{rel(gc, gp):.2e} seeds {rel(gs, gsp):.2e}; vs FP64 trajectory (lines agreeing "
          f"{ok.sum()}/{len(ok)}): charges {rel(gc, gc64):.2e}", flush=True)
    assert rel(gc, gp) < 1e-4 and rel(gs, gsp) < 1e-4
    assert ok.mean() > 0.8
    assert rel(gc, gc64) < 1e-2


def test_checkpoint_interval_does_not_change_result(gpu):
    """Segment recompute is bit-identical to the forward, so any interval gives the same bits."""
    w = gen.plain_config(300, 200, 6.0, 90, seed=5)
    G = np.random.default_rng(1).normal(size=(3, w.len, w.n))
    ref = None
    for every in (1, 7, 50, 89, 200):
        _, gc, gs = forward_and_adjoint(gpu, w.charges, w.seeds, w.len, rules.angle(), G, every)
        if ref is None:
            ref = (gc, gs)
        else:
            assert np.array_equal(gc, ref[0]) and np.array_equal(gs, ref[1]), every


def test_repeated_calls_reuse_the_work_lists(gpu):
    """The single-configuration adjoint keeps its work lists between calls:
    repeated calls, calls alternating between two problems (which rebuild
    them), a call with new gradients on the same problem, and a call on a
    second stream all give the bits of a first call."""
    wa = gen.plain_config(300, 200, 6.0, 90, seed=5)
    wb = gen.plain_config(120, 96, 4.0, 50, seed=6)
    rng = np.random.default_rng(3)
    Ga, Ga2 = rng.normal(size=(3, wa.len, wa.n)), rng.normal(size=(3, wa.len, wa.n))
    Gb = rng.normal(size=(3, wb.len, wb.n))
    rule = rules.angle()
    first_a = forward_and_adjoint(gpu, wa.charges, wa.seeds, wa.len, rule, Ga, 7)
    first_a2 = forward_and_adjoint(gpu, wa.charges, wa.seeds, wa.len, rule, Ga2, 7)
    first_b = forward_and_adjoint(gpu, wb.charges, wb.seeds, wb.len, rule, Gb, 7)
    for _ in range(2):
        for w, G, want in ((wa, Ga, first_a), (wa, Ga2, first_a2), (wa, Ga, first_a),
                           (wb, Gb, first_b)):
            got = forward_and_adjoint(gpu, w.charges, w.seeds, w.len, rule, G, 7)
            assert all(np.array_equal(x, y) for x, y in zip(got, want))
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        got = forward_and_adjoint(gpu, wa.charges, wa.seeds, wa.len, rule, Ga, 7)
    torch.cuda.synchronize()
    assert all(np.array_equal(x, y) for x, y in zip(got, first_a))


def test_adjoint_multi_tile_matches_single_tile(gpu):
    w = gen.plain_config(2000, 128, 12.0, 40, seed=8)
    G = np.random.default_rng(2).normal(size=(3, w.len, w.n)).astype(np.float32)
    out = []
    for tile in (0, 512):
        c = dev(w.charges)
        rows = gpu.seed_rows(dev(w.seeds), w.len)
        launch = gpu.make_launch(2, tile, 256)
        ck = torch.empty(gpu.ckpt_size(w.n, w.len, 10), dtype=torch.float32, device="cuda")
        gpu.trace_into(c, rows, w.n, rules.angle(), launch, ck, 10)
        gc, _ = gpu.adjoint(c, rows, ck, 10, dev(G), rules.angle(), launch)
        out.append(gc.cpu().numpy().astype(np.float64))
    assert rel(out[1], out[0]) < 1e-3


def _adjoint_2k_problem(seed=0, sigma=0.05):
    w = gen.adjoint_2k(seed)
    target_q = gen.perturbed(w.charges, sigma, seed=seed)
    return w, target_q


def agreeing_lines(g_rows, g_steps, r64, s64, tol=1e-3):
    """Lines whose float32 GPU trajectory follows the FP64 one: same step
    count and every point within tol.  Elsewhere the two traces took different
    discrete decisions (where a line stops, which clamp applies), and their
    gradients describe different piecewise-smooth branches of the loss."""
    err = np.abs(g_rows.astype(np.float64) - r64).max(axis=(0, 1))
    return (g_steps == s64) & (err <= tol)


def test_adjoint_2k_gradient_vs_fp64_adjoint_and_fd(gpu, orc):
    """256 charges, 2,048 lines, 500 rows, angle rule with stop, checkpoints
    every 50 steps; loss = mean |x - x_target|^2."""
    rule = rules.angle()
    w, tq = _adjoint_2k_problem()
    npts = w.n * w.len
    tgt = gpu.trace(dev(tq), dev(w.seeds), w.len, rule)
    # full autograd gradient (reported)
    c = dev(w.charges).requires_grad_(True)
    rows = gpu.trace_lines(c, dev(w.seeds), w.len, rule, 50)
    loss = ((rows - tgt) ** 2).sum() / npts
    loss.backward()
    g_full = c.grad.cpu().numpy().astype(np.float64)
    # per-line agreement with FP64
    g_steps = forward_steps(gpu, w.charges, w.seeds, w.len, rule, 50)
    r64, s64 = orc.trace64(w.charges, w.seeds, w.len, rule, with_steps=True)
    g_rows = rows.detach().cpu().numpy()
    ok = agreeing_lines(g_rows, g_steps, rows64_to_soa(r64), s64.astype(np.int64))
    G = (2.0 / npts) * (rows.detach() - tgt).cpu().numpy().astype(np.float64)
    G[:, :, ~ok] = 0.0
    # masked gradients: GPU adjoint and FP64 adjoint of the same loss
    n = w.n
    cc = dev(w.charges)
    rr = gpu.seed_rows(dev(w.seeds), w.len)
    ck = torch.empty(gpu.ckpt_size(n, w.len, 50), dtype=torch.float32, device="cuda")
    gpu.trace_into(cc, rr, n, rule, None, ck, 50)
    gm, _ = gpu.adjoint(cc, rr, ck, 50, dev(G.astype(np.float32)), rule)
    torch.cuda.synchronize()
    g_gpu = gm.cpu().numpy().astype(np.float64)
    G64 = np.ascontiguousarray(G.transpose(1, 2, 0))
    g64, _ = orc.adjoint64(w.charges, w.seeds, w.len, rule, G64)
    g64_full, _ = orc.adjoint64(w.charges, w.seeds, w.len, rule,
                                np.ascontiguousarray(((2.0 / npts) * (rows.detach() - tgt).cpu()
                                                      .numpy().astype(np.float64)).transpose(1, 2, 0)))
    r_adj = rel(g_gpu, g64)
    # the same dL/dx, FP64 sweep along the GPU's own path (all lines)
    Gf = np.ascontiguousarray(((2.0 / npts) * (rows.detach() - tgt).cpu().numpy()
                               .astype(np.float64)).transpose(1, 2, 0))
    g_path, _ = orc.adjoint64_path(w.charges, w.seeds, w.len, rule, Gf, g_rows, g_steps)
    r_path = rel(g_full, g_path)
    # FP64 central differences of the FP64 loss, on lines whose discrete
    # decisions do not move: the lines that agree with the GPU trace and whose
    # step count stays the same under every +-h perturbation below (the stop
    # rule makes the loss jump where a line ends one step earlier or later).
    # 24 parameters: the 12 largest FP64 gradient components and 12 at random.
    tgt64 = orc.trace64(tq, w.seeds, w.len, rule)
    sel = np.nonzero(ok)[0]
    base = w.charges.astype(np.float64)
    G64m = np.zeros_like(r64)
    G64m[:, sel, :] = (2.0 / npts) * (r64[:, sel, :] - tgt64[:, sel, :])
    g64m, _ = orc.adjoint64(w.charges, w.seeds, w.len, rule, G64m)
    rng = np.random.default_rng(7)
    picks = np.unique(np.concatenate([np.argsort(-np.abs(g64m).ravel())[:12],
                                      rng.choice(g64m.size, 12, replace=False)]))

    def per_line(ch):
        r, st = orc.trace64(ch, w.seeds[sel], w.len, rule, with_steps=True)
        return ((r - tgt64[:, sel, :]) ** 2).sum(axis=(0, 2)) / npts, st.astype(np.int64)

    _, base_steps = per_line(base)
    stable = np.ones(sel.size, bool)
    diffs = {}
    for h in (1e-4, 1e-5):
        rowsd = []
        for flat in picks:
            i, k = divmod(int(flat), 4)
            cp, cm = base.copy(), base.copy()
            cp[i, k] += h
            cm[i, k] -= h
            lp, sp = per_line(cp)
            lm, sm = per_line(cm)
            stable &= (sp == base_steps) & (sm == base_steps)
            rowsd.append((lp - lm) / (2 * h))
        diffs[h] = np.array(rowsd)                    # (params, lines)
    keep = sel[stable]
    fd = {h: diffs[h][:, stable].sum(axis=1) for h in diffs}
    # parameters where the h = 1e-4 quotient has converged (no clamp kink inside +-h)
    smooth = np.abs(fd[1e-4] - fd[1e-5]) <= 1e-2 * np.abs(fd[1e-5]) + 1e-12
    # the FP64 adjoint and the GPU adjoint of that same loss (lines `keep`)
    G64k = np.zeros_like(r64)
    G64k[:, keep, :] = (2.0 / npts) * (r64[:, keep, :] - tgt64[:, keep, :])
    g64k, _ = orc.adjoint64(w.charges, w.seeds, w.len, rule, G64k)
    Gk = np.zeros_like(G)
    Gk[:, :, keep] = ((2.0 / npts) * (rows.detach() - tgt).cpu().numpy().astype(np.float64))[:, :, keep]
    gk, _ = gpu.adjoint(cc, rr, ck, 50, dev(Gk.astype(np.float32)), rule)
    torch.cuda.synchronize()
    g_gpu_k = gk.cpu().numpy().astype(np.float64)
    f = fd[1e-4][smooth]
    r_fd64 = float(np.linalg.norm(g64k.ravel()[picks][smooth] - f) / np.linalg.norm(f))
    r_fd = float(np.linalg.norm(g_gpu_k.ravel()[picks][smooth] - f) / np.linalg.norm(f))
    worst = float(np.max(np.abs(g64k.ravel()[picks][smooth] - f) / np.abs(f)))
    print(f"[adjoint_2k] loss {loss.item():.4e}; agreeing lines {ok.sum()}/{len(ok)}; "
          f"mean steps {g_steps.mean():.1f}; rel(gpu, FP64 along gpu path) {r_path:.2e}; "
          f"(a) rel(gpu, fp64 adjoint), same dL/dx: {r_adj:.2e} "
          f"[all lines: {rel(g_full, g64_full):.2e}]; FD lines with fixed step counts "
          f"{keep.size}/{sel.size}; FD(1e-4)-converged params {smooth.sum()}/{len(picks)}: "
          f"(b) rel(fp64 adjoint, FD) {r_fd64:.2e} (worst component {worst:.2e}), "
          f"(c) rel(gpu, FD) {r_fd:.2e}", flush=True)
    assert r_path < 1e-3
    assert ok.mean() > 0.8
    assert r_adj < 5e-2
    assert keep.size >= 0.5 * sel.size
    assert smooth.sum() >= 10
    assert r_fd64 < 1e-2
    assert r_fd < 5e-2


def test_autograd_seed_gradient_and_grad_of_row0(gpu):
    w = gen.plain_config(64, 64, 4.0, 30, seed=2)
    s = dev(w.seeds).requires_grad_(True)
    c = dev(w.charges).requires_grad_(True)
    rows = gpu.trace_lines(c, s, w.len, rules.plain(10.0), 5)
    (rows[:, 0, :].sum() * 0 + rows.sum()).backward()
    assert s.grad.shape == (w.n, 3) and torch.isfinite(s.grad).all()
    assert c.grad.shape == (w.p, 4) and torch.isfinite(c.grad).all()


def test_ref_rule_adjoint_is_refused(gpu):
    w = gen.plain_config(16, 64, 4.0, 10, seed=2)
    c = dev(w.charges)
    rows = gpu.seed_rows(dev(w.seeds), w.len)
    ck = torch.empty(gpu.ckpt_size(w.n, w.len, 5), dtype=torch.float32, device="cuda")
    gpu.trace_into(c, rows, w.n, rules.ref(10.0), None, ck, 5)
    with pytest.raises(gpu.EmError) as e:
        gpu.adjoint(c, rows, ck, 5, torch.zeros_like(rows), rules.ref(10.0))
    assert e.value.code == 20
