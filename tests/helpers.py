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
