"""GPU parity tests of the forward tracer (HIP path through the C ABI).

Parity bar (float32 trajectories), applied line by line (tests/helpers.py,
``line_gates``): over each line's prefix -- before its first approach within
0.05 of a charge, the whole line for a line that stays farther -- the GPU
points lie within 1e-4 of the CPU path (pointwise for the plain and ref
rules; distance to the CPU path's polyline for the angle rule, whose adaptive
step makes the along-line position an ill-conditioned function of rounding),
unless that line is shown to be beyond float32 resolution -- the CPU float32
path departs from FP64 by more than 2e-5 there, or the FP64 path itself moves
by more than 2e-5 under float32-sized perturbations -- and the GPU-CPU gap is
within 10x that departure or move.  Each exemption is printed with its
numbers.  Over all lines the GPU's distance to FP64 is no
worse than the CPU float32 path's (median and 90th percentile, 2x).
Step counts: len - 1 on every line without the stop rule; with it, identical
to the CPU path's or decided by rounding at the deciding step (``step_gate``).
Field evaluations: within 1e-5 relative of the FP64 field where the field is
not dominated by cancellation (|E| >= 0.1 * sum_i |E_i|).
"""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from emtrace import gen, rules
from helpers import (COND, TOL, line_errors, line_gates, parity_report, prefix_lengths,
                     rows64_to_soa, step_gate)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_trace(em, w_charges, seeds, length, rule, launch=None):
    rows = em.trace(dev(w_charges), dev(seeds), length, rule, launch)
    torch.cuda.synchronize()
    return rows.cpu().numpy()


def print_report(tag, rep):
    keys = ["lines", "eligible_lines", "eligible_within_tol", "prefix_within_tol",
            "prefix_median_err", "prefix_max_err", "path_within_tol", "path_max_err"]
    print(f"[parity] {tag}: " + ", ".join(f"{k}={rep[k]:.3g}" if isinstance(rep[k], float)
                                         else f"{k}={rep[k]}" for k in keys), flush=True)


def _golden_f32():
    return sorted(p for p in os.listdir(GOLDEN) if p.startswith("ref_f32"))


@pytest.mark.parametrize("name", _golden_f32())
def test_gpu_matches_reference_golden(gpu, orc, name):
    """Reference CalcField_CPU<float> output (compiled reference, golden
    vectors) vs the HIP path, line by line.  The tolerances scale with the
    step length (1e-4 and 2e-5 at step 0.1): over the same number of rows a
    longer step carries the direction's rounding proportionally farther."""
    g = np.load(os.path.join(GOLDEN, name))
    res = float(g["resolution"])
    rule = rules.ref(res) if int(g["use_curvature"]) else rules.plain(res)
    length = g["rows"].shape[1]
    rows, steps = gpu.trace(dev(g["charges"]), dev(g["seeds"]), length, rule, with_steps=True)
    rows = rows.cpu().numpy()
    assert np.array_equal(rows[:, 0], g["rows"][:, 0])
    assert (steps.cpu().numpy() == length - 1).all()
    rc, r64 = orc.calc_lines(g["charges"], g["seeds"], length, rules.plain(res), dtype=np.float64)
    assert rc == 0
    k = (1.0 / res) / 0.1
    line_gates(name, rows, g["rows"], r64, g["charges"], g["seeds"], rule, tol=TOL * k,
               cond=COND * k, min_conditioned=0.8)


_FP64_ROWS = {}


def _lines_8k_fp64(orc, w, rule):
    """FP64 rows of all lines_8k lines (the restatement in double for plain
    and ref -- (k+1) == 1 in double as in float -- the FP64 trajectory for the
    angle rule), computed once per rule."""
    key = (rule.rule == rules.ANGLE, rule.stop)
    if key not in _FP64_ROWS:
        if rule.rule == rules.ANGLE:
            _FP64_ROWS[key] = rows64_to_soa(orc.trace64(w.charges, w.seeds, w.len, rule))
        else:
            rc, r64 = orc.calc_lines(w.charges, w.seeds, w.len, rules.plain(rule.resolution),
                                     dtype=np.float64)
            assert rc == 0
            _FP64_ROWS[key] = r64
    return _FP64_ROWS[key]


@pytest.mark.parametrize("rule", [rules.ref(10.0), rules.plain(10.0)], ids=["ref", "plain"])
def test_full_lines_8k_vs_compiled_reference(gpu, orc, rule):
    """All of lines_8k (1,000 charges, 8,192 lines, 2,500 rows, resolution 10)
    against the reference's own CalcField_CPU<float> (CPU_Implement.cpp:
    207-351 with useCurvature, 38-102 without), compiled from source: every
    line takes len - 1 steps, and every line passes the per-line gate."""
    if oracle.ref_lib() is None:
        pytest.skip("compiled reference absent (oracle/_ref)")
    w = gen.lines_8k()
    rows_t, steps_t = gpu.trace(dev(w.charges), dev(w.seeds), w.len, rule, with_steps=True)
    torch.cuda.synchronize()
    g_rows, g_steps = rows_t.cpu().numpy(), steps_t.cpu().numpy()
    del rows_t
    curv = int(rule.rule == rules.REF)
    rc, c_rows, t_ref, _ = orc.ref_calc_lines(w.charges, w.seeds, w.len, rule.resolution, curv)
    assert rc == 0
    print(f"[full lines_8k] {rules.NAMES[rule.rule]}: compiled reference {t_ref:.2f} s", flush=True)
    assert (g_steps == w.len - 1).all()
    r64 = _lines_8k_fp64(orc, w, rule)
    line_gates(f"lines_8k all/{rules.NAMES[rule.rule]} vs reference", g_rows, c_rows, r64,
               w.charges, w.seeds, rule, min_conditioned=0.85)


@pytest.mark.parametrize("rule", [rules.angle(), rules.angle(stop=False)],
                         ids=["angle", "angle-nostop"])
def test_full_lines_8k_angle_vs_cpu_path(gpu, orc, rule):
    """All of lines_8k with the angle rule (no reference counterpart) against
    the C restatement of the CPU path in the HIP code's operation order: step
    counts per line (step_gate), points per line (path distance)."""
    w = gen.lines_8k()
    rows_t, steps_t = gpu.trace(dev(w.charges), dev(w.seeds), w.len, rule, with_steps=True)
    torch.cuda.synchronize()
    g_rows, g_steps = rows_t.cpu().numpy(), steps_t.cpu().numpy().astype(np.int64)
    del rows_t
    rc, c_rows, c_steps = orc.calc_lines(w.charges, w.seeds, w.len, rule, with_steps=True)
    assert rc == 0
    tag = f"lines_8k all/{rules.NAMES[rule.rule]}{'' if rule.stop else '-nostop'}"
    if rule.stop:
        print(f"[steps] {tag}: mean steps {g_steps.mean():.1f}", flush=True)
        step_gate(tag, w.charges, g_rows, g_steps, c_rows, c_steps)
    else:
        assert (g_steps == w.len - 1).all() and (c_steps == w.len - 1).all()
    r64 = _lines_8k_fp64(orc, w, rule)
    line_gates(tag, g_rows, c_rows, r64, w.charges, w.seeds, rule, min_conditioned=0.5)


def test_field_matches_fp64(gpu, orc):
    """Per-evaluation field vs FP64: <= 1e-5 relative where not cancellation-dominated."""
    w = gen.lines_8k()
    rng = np.random.default_rng(3)
    pts = np.concatenate([rng.uniform(-12, 12, (4000, 3)),
                          w.seeds[:2000].astype(np.float64)]).astype(np.float32)
    for scale in (1.0, rules.K_E):
        E = gpu.field(dev(w.charges), dev(pts), scale).cpu().numpy().astype(np.float64)
        E64 = orc.field64(w.charges.astype(np.float64), pts.astype(np.float64), scale)
        d = pts.astype(np.float64)[:, None, :] - w.charges[None, :, :3].astype(np.float64)
        r = np.linalg.norm(d, axis=2)
        mag = (np.abs(w.charges[None, :, 3]) / r ** 2).sum(axis=1) * scale
        n64 = np.linalg.norm(E64, axis=1)
        rel = np.linalg.norm(E - E64, axis=1) / n64
        ok = n64 >= 0.1 * mag
        print(f"[field] scale={scale:.3g}: {ok.sum()} well-conditioned points, max rel "
              f"{rel[ok].max():.3g}, median {np.median(rel[ok]):.3g}; all points max "
              f"|dE|/sum|E_i| {(np.linalg.norm(E - E64, axis=1) / mag).max():.3g}", flush=True)
        assert ok.sum() > 1000
        assert rel[ok].max() <= 1e-5
        assert (np.linalg.norm(E - E64, axis=1) / mag).max() <= 1e-6


def test_single_charge_radial_lines(gpu):
    c = np.array([[1.25, -0.5, 2.0, 0.8]], np.float32)
    d = gen.fibonacci_dirs(256)
    seeds = (c[0, :3].astype(np.float64) + 0.1 * d).astype(np.float32)
    rows = gpu_trace(gpu, c, seeds, 60, rules.plain(10.0))
    s = np.arange(60)[:, None, None]
    expect = c[0, :3][None, None, :] + (0.1 + s * 0.1) * d[None]
    assert np.abs(rows.transpose(1, 2, 0) - expect).max() < 2e-5
    rows = gpu_trace(gpu, c, seeds, 30, rules.angle())
    r = np.array([0.1, 0.2] + [0.2 + 0.5 * k for k in range(1, 29)])
    expect = c[0, :3][None, None, :] + r[:, None, None] * d[None]
    assert np.abs(rows.transpose(1, 2, 0) - expect).max() < 5e-5


def test_ref_rule_bitwise_equals_plain(gpu):
    """(k+1) rounds to 1 with k_e in E, so the reference's curvature step is its plain step."""
    w = gen.plain_config(500, 1024, 10.0, 300, seed=1)
    a = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.plain(10.0))
    b = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.ref(10.0))
    same = (a == b).all(axis=(0, 1))
    print(f"[ref==plain] identical lines {same.sum()}/{len(same)}", flush=True)
    # on every line, up to its first approach within 0.05 of a charge (closer,
    # |(E-P) x P|^2 can overflow float and the ref rule's k is no longer 0)
    first, _ = prefix_lengths(a, w.charges)
    for l in range(w.n):
        assert np.array_equal(a[:, :first[l], l], b[:, :first[l], l]), l


@pytest.mark.parametrize("S", [1, 2, 4, 8, 16, 32])
def test_lanes_per_line_agree(gpu, S):
    w = gen.plain_config(700, 300, 10.0, 200, seed=6)
    base = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.plain(10.0),
                     launch=gpu.make_launch(1, 0, 256))
    rows = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.plain(10.0),
                     launch=gpu.make_launch(S, 0, 256))
    rep, *_ = parity_report(rows, base, w.charges)
    assert rep["prefix_median_err"] < 1e-5 and rep["prefix_within_tol"] > 0.95


def test_multi_tile_agrees_with_single_tile(gpu):
    w = gen.plain_config(3000, 256, 12.0, 120, seed=8)
    single = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.angle(), launch=gpu.make_launch(2, 0, 256))
    multi = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.angle(), launch=gpu.make_launch(2, 512, 256))
    rep, *_ = parity_report(multi, single, w.charges)
    assert rep["path_within_tol"] > 0.95


def test_deterministic_and_scale_invariant(gpu):
    w = gen.plain_config(800, 512, 10.0, 150, seed=12)
    a = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.angle())
    b = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.angle())
    assert np.array_equal(a, b)
    q2 = w.charges.copy()
    q2[:, 3] *= 2.0  # exact in float: every partial sum doubles, directions are bitwise equal
    c = gpu_trace(gpu, q2, w.seeds, w.len, rules.angle())
    assert np.array_equal(a, c)


def test_drop_in_error_codes(gpu):
    L = gpu.lib()
    ch = np.zeros((4, 4), np.float32)
    x = np.zeros(640, np.float32)
    y = np.zeros(640, np.float32)
    z = np.zeros(640, np.float32)
    P = lambda a: a.ctypes.data  # noqa: E731
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 0, 10, 10.0, 1, None) == 1
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, 0.0, 1, None) == 2
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 64, 1, 10.0, 1, None) == 3
    # unlike the SSE path (status 5), n need not be a multiple of 64
    w = gen.plain_config(50, 40, 5.0, 16, seed=2)
    rows = np.zeros((3, w.len, w.n), np.float32)
    rows[:, 0] = w.seeds.T
    rc, perf = gpu.calc_lines(w.charges, rows[0], rows[1], rows[2], w.n, 10.0, True)
    assert rc == 0 and perf["progress"] == 1.0 and perf["kernel_s"] > 0
    assert np.isfinite(rows).all()
    assert perf["interactions"] == 50 * 40 * 15


def test_drop_in_matches_device_path(gpu):
    w = gen.plain_config(1000, 640, 10.0, 100, seed=4)
    rows = np.zeros((3, w.len, w.n), np.float32)
    rows[:, 0] = w.seeds.T
    rc, perf = gpu.calc_lines(w.charges, rows[0], rows[1], rows[2], w.n, 10.0, True)
    assert rc == 0
    dev_rows = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.ref(10.0))
    assert np.array_equal(rows, dev_rows)
    assert perf["h2d_bytes"] > 0 and perf["d2h_bytes"] == 3 * 4 * w.n * (w.len - 1)


EDGE = [
    # (name, p, n, len, rule, launch (S, tile, block) or None)
    ("one_line", 50, 1, 200, rules.plain(10.0), None),
    ("two_rows", 300, 130, 2, rules.angle(), None),
    ("ragged_65_lines", 100, 65, 120, rules.ref(10.0), None),
    ("p_equals_tile", 256, 96, 80, rules.plain(10.0), (2, 256, 256)),
    ("p_tile_plus_one", 257, 96, 80, rules.plain(10.0), (2, 256, 256)),
    ("p_tile_minus_one", 255, 96, 80, rules.angle(), (2, 256, 256)),
    ("wide_groups_tiny_p", 40, 64, 60, rules.angle(), (32, 512, 64)),
    # the largest charge sets one LDS tile holds beside the kernels' static LDS
    # (9,984 records + 512 B within 160 KiB), and just past it
    ("p_largest_single_tile", 9984, 64, 40, rules.plain(10.0), None),
    ("p_past_single_tile", 10000, 64, 40, rules.angle(stop=False), None),
    ("p_10240", 10240, 64, 40, rules.plain(10.0), None),
]


@pytest.mark.parametrize("case", EDGE, ids=[c[0] for c in EDGE])
def test_edge_shapes_vs_cpu_path(gpu, orc, case):
    """Shapes at the edges of the work split and of the charge tiles, against
    the CPU path: the same step count on every line, every line through the
    per-line gate (REF needs n % 64 on the CPU path: lines are padded there
    and cut)."""
    name, p, n, length, rule, shape = case
    w = gen.plain_config(p, n, 5.0, length, seed=21)
    launch = gpu.make_launch(shape[0], shape[1], shape[2]) if shape else None
    rows, steps = gpu.trace(dev(w.charges), dev(w.seeds), length, rule, launch, with_steps=True)
    rows = rows.cpu().numpy()
    seeds = w.seeds
    if rule.rule == rules.REF and n % 64:
        seeds = np.concatenate([seeds, np.repeat(seeds[:1], 64 - n % 64, axis=0)])
    rc, c_rows, c_steps = orc.calc_lines(w.charges, seeds, length, rule, with_steps=True)
    assert rc == 0
    c_rows, c_steps = c_rows[:, :, :n], c_steps[:n]
    assert np.array_equal(steps.cpu().numpy().astype(np.int64), c_steps.astype(np.int64))
    if rule.rule == rules.ANGLE:
        r64 = rows64_to_soa(orc.trace64(w.charges, w.seeds, length, rule))
    else:
        rc, r64 = orc.calc_lines(w.charges, w.seeds, length, rules.plain(rule.resolution),
                                 dtype=np.float64)
        assert rc == 0
    first, _ = prefix_lengths(c_rows, w.charges)
    err = line_errors(rows, c_rows, first)
    print(f"[edge {name}] lines {n}: median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) <= 1e-5
    line_gates(f"edge {name}", rows, c_rows, r64, w.charges, w.seeds, rule, min_conditioned=0.3)


def test_tile_beyond_the_lds_budget_is_rejected(gpu):
    """A tile must leave room for the kernels' static LDS: 10,240 records
    (exactly 160 KiB) is refused before any launch; 9,984 runs."""
    from emtrace._lib import EmError
    w = gen.plain_config(10240, 64, 5.0, 8, seed=3)
    with pytest.raises(EmError) as e:
        gpu_trace(gpu, w.charges, w.seeds, w.len, rules.plain(10.0), gpu.make_launch(1, 10240, 256))
    assert "static LDS" in str(e.value)
    rows = gpu_trace(gpu, w.charges, w.seeds, w.len, rules.plain(10.0), gpu.make_launch(1, 9984, 256))
    assert np.isfinite(rows).all()


def test_no_charges_gives_the_reference_nan(gpu, orc):
    """p = 0: E = 0 and the plain step divides 0 by 0, as the reference does."""
    w = gen.plain_config(4, 64, 5.0, 6, seed=2)
    empty = np.zeros((0, 4), np.float32)
    rows = gpu.trace(dev(empty), dev(w.seeds), 6, rules.plain(10.0)).cpu().numpy()
    rc, c_rows = orc.calc_lines(empty, w.seeds, 6, rules.plain(10.0))
    assert rc == 0
    assert np.array_equal(rows[:, 0], c_rows[:, 0])
    assert np.isnan(rows[:, 1:]).all() and np.isnan(c_rows[:, 1:]).all()


@pytest.mark.parametrize("rule", [rules.ref(1.0), rules.plain(1.0)], ids=["ref", "plain"])
def test_host_path_row_chunks_match_device_path(gpu, rule):
    """The reference's default problem (1,024 charges, 16,384 lines, 2,500
    rows) through the host-array entry point: its 491 MB of output is traced
    in row chunks whose copies back overlap the next chunk's kernel.  Each
    chunk continues from the previous chunk's last row, so the rows and step
    counts are bit-identical to one device-side launch."""
    w = gpu.legacy.preset("default")
    rows_h, perf, steps_h = gpu.trace_host(w.charges, w.seeds, w.len, rule)
    rows_d, steps_d = gpu.trace(dev(w.charges), dev(w.seeds), w.len, rule, with_steps=True)
    torch.cuda.synchronize()
    assert perf["progress"] == 1.0 and perf["kernel_s"] > 0
    assert np.array_equal(steps_h.astype(np.int64), steps_d.cpu().numpy().astype(np.int64))
    assert np.array_equal(rows_h, rows_d.cpu().numpy())


def test_drop_in_mirrors_progress(gpu):
    """em_perf.progress_out: the host-array call stores every progress update
    there as well (the drop-in points it at perfPacket.progress, which the
    reference's monitor thread polls); the last value is 1.0."""
    from emtrace import _lib
    w = gpu.legacy.preset("default")            # 491 MB of rows: 8 row chunks
    x, y, z = (np.empty(w.len * w.n, np.float32) for _ in range(3))
    x[:w.n], y[:w.n], z[:w.n] = w.seeds[:, 0], w.seeds[:, 1], w.seeds[:, 2]
    seen = ctypes.c_double(-1.0)
    perf = _lib.em_perf()
    perf.progress_out = ctypes.pointer(seen)
    P = lambda a: a.ctypes.data  # noqa: E731
    rc = gpu.lib().em_calc_lines_f32(P(w.charges), w.p, P(x), P(y), P(z), w.n, w.len, 1.0, 1,
                                     ctypes.byref(perf))
    assert rc == 0
    assert seen.value == 1.0 and perf.progress == 1.0
