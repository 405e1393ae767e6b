"""FP64 kernel (the CalcField_CPU<double> counterpart).

* every row equals the oracle's FP64 restatement bit for bit -- plain, ref,
  angle with and without line ends, single- and multi-tile charge staging,
  partial workgroups, one, two and four lanes per line, pitch > n on the
  device form;
* the square roots and divisions that skip the compiler's range wrappers
  inside the fast domain give the same bits as the compiler's expansions
  (all of lines_64k; four-lane groups, each lane adding one component, on the
  reference's default problem and on 16,384 lines_64k lines) and as the
  restatement at the domain's edges;
* the host-array form traces the plain and ref rules in row chunks, bit-equal
  to one device launch;
* against the reference's own double rows (golden vectors) it is as close as
  the restatement is (-ffast-math reassociation in the reference build);
* the drop-in calc_lines64 keeps the reference's status codes;
* at the full lines_8k size (1,000 charges, 8,192 lines, 2,500 rows) the FP32
  kernel drifts from the FP64 kernel no more than the CPU FP32 path does;
* the command line with --cpuprecision=double --gpuprecision=double writes
  identical CPU and GPU text.
"""
import dataclasses
import glob
import os

import numpy as np
import pytest
import torch

from emtrace import gen, legacy, rules

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _as_f32_params(rule):
    """The em_step carries float32 parameters: give the oracle the same values."""
    f = lambda v: float(np.float32(v))  # noqa: E731
    return dataclasses.replace(rule, resolution=f(rule.resolution), h0=f(rule.h0), hmin=f(rule.hmin),
                               hmax=f(rule.hmax), turn=f(rule.turn))


CASES = [
    # (name, p, n, half_side, len, rule)
    ("plain", 300, 512, 10.0, 200, rules.plain(10.0)),
    ("ref", 300, 512, 10.0, 200, rules.ref(10.0)),
    ("angle_stop", 300, 448, 5.0, 300, rules.angle()),
    ("angle_nostop", 300, 448, 5.0, 300, rules.angle(stop=False)),
    ("ragged_block", 7, 130, 3.0, 50, rules.angle()),
    ("multi_tile_plain", 5000, 128, 20.0, 30, rules.plain(10.0)),
    ("multi_tile_angle", 4100, 192, 20.0, 40, rules.angle()),
    # 32,768 < n lines: one lane per line (the cases above take four); two lanes
    # per line, only by EMTRACE_FP64_LANES; 67 charges leave a last pass of one term
    ("one_lane_plain", 67, 40000, 10.0, 30, rules.plain(10.0)),
    ("two_lanes_plain", 67, 40000, 10.0, 30, rules.plain(10.0)),
    ("two_lanes_angle", 67, 40000, 5.0, 40, rules.angle()),
    ("four_lanes_angle", 67, 40000, 5.0, 40, rules.angle()),
    # eight lanes per line (groups across two quads of lanes), single and staged tiles
    ("eight_lanes_ref", 300, 512, 10.0, 200, rules.ref(10.0)),
    ("eight_lanes_angle", 300, 448, 5.0, 300, rules.angle()),
    ("eight_lanes_multi_tile_plain", 5000, 128, 20.0, 30, rules.plain(10.0)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_bitwise_equals_restatement(gpu, orc, case, monkeypatch):
    name, p, n, half, length, rule = case
    if name.startswith("two_lanes"):
        monkeypatch.setenv("EMTRACE_FP64_LANES", "2")
    elif name.startswith("four_lanes"):
        monkeypatch.setenv("EMTRACE_FP64_LANES", "4")
    elif name.startswith("eight_lanes"):
        monkeypatch.setenv("EMTRACE_FP64_LANES", "8")
    rule = _as_f32_params(rule)
    w = gen.plain_config(p, n, half, length, seed=11)
    ch = w.charges.astype(np.float64)
    sd = w.seeds.astype(np.float64)
    rc, want, want_steps = orc.calc_lines(ch, sd, length, rule, dtype=np.float64, with_steps=True)
    assert rc == 0
    rows, steps = gpu.trace64(torch.from_numpy(ch).cuda(), torch.from_numpy(sd).cuda(), length, rule,
                              with_steps=True)
    rows = rows.cpu().numpy()
    same = (rows == want) | (np.isnan(rows) & np.isnan(want))
    print(f"[fp64 {name}] identical {same.mean():.6f}; steps identical "
          f"{np.mean(steps.cpu().numpy() == want_steps.astype(np.int64)):.4f}", flush=True)
    assert same.all()
    assert np.array_equal(steps.cpu().numpy(), want_steps.astype(np.int32))


def _edge_problem(kind, p, n, length, seed):
    """Problems at the edges of the kernel's fast square-root and division
    domain (coordinates 0 or of magnitude 2^-100 .. 2^190, |q| 2^-100 .. 2^100)."""
    w = gen.plain_config(p, n, 5.0, length, seed=seed)
    ch, sd = w.charges.astype(np.float64), w.seeds.astype(np.float64)
    if kind == "planar":  # r_z = 0: zero numerators inside the domain
        ch[:, 2] = 0.0
        sd[:, 2] = 0.0
    elif kind == "tiny_coordinate":  # 1e-35 < 2^-100: that charge's tile leaves the domain
        ch[p // 2, 0] = 1e-35
    elif kind == "zero_charge":
        ch[p // 3, 3] = 0.0
    elif kind == "far_charge":  # 1e60 > 2^190
        ch[1, :3] = (1e60, -2e60, 3e60)
    elif kind == "tiny_point":  # one line's first point leaves the domain
        sd[5, 1] = 1e-40
    else:
        raise ValueError(kind)
    return ch, sd


EDGES = [(kind, p, n, length) for kind in ("planar", "tiny_coordinate", "zero_charge", "far_charge",
                                           "tiny_point")
         for (p, n, length) in ((300, 512, 60), (3000, 128, 12))] + [
    # one lane per line (65,536 lines)
    ("planar", 40, 65536, 8), ("tiny_coordinate", 40, 65536, 8)]


@pytest.mark.parametrize("edge", EDGES, ids=[f"{e[0]}-p{e[1]}-n{e[2]}" for e in EDGES])
@pytest.mark.parametrize("rule", [rules.plain(10.0), rules.ref(10.0), rules.angle()],
                         ids=["plain", "ref", "angle"])
def test_fp64_fast_domain_edges_bitwise(gpu, orc, edge, rule):
    """Inside the fast domain the kernel's square roots and divisions skip the
    compiler's range wrappers; outside it (a tile with a charge beyond it, a
    wave with a point beyond it) the compiler's expansions run.  Both give the
    restatement's bits."""
    kind, p, n, length = edge
    rule = _as_f32_params(rule)
    ch, sd = _edge_problem(kind, p, n, length, seed=19)
    rc, want, want_steps = orc.calc_lines(ch, sd, length, rule, dtype=np.float64, with_steps=True)
    assert rc == 0
    rows, steps = gpu.trace64(torch.from_numpy(ch).cuda(), torch.from_numpy(sd).cuda(), length, rule,
                              with_steps=True)
    rows = rows.cpu().numpy()
    same = (rows == want) | (np.isnan(rows) & np.isnan(want))
    print(f"[fp64 edge {kind} p={p} n={n} {rules.NAMES[rule.rule]}] identical {same.mean():.6f}",
          flush=True)
    assert same.all()
    assert np.array_equal(steps.cpu().numpy(), want_steps.astype(np.int32))


@pytest.mark.parametrize("rule", [rules.plain(10.0), rules.ref(10.0), rules.angle()],
                         ids=["plain", "ref", "angle"])
def test_fp64_fast_form_equals_compiler_expansions_full_lines_64k(gpu, monkeypatch, rule):
    """All 16,384 lines_64k charges and 65,536 lines (one lane per line, staged
    tiles), 40 rows: the fast square roots and divisions against the compiler's
    expansions (EMTRACE_FP64_EXACT=1), bit for bit over 4.2e10 terms."""
    w = gen.lines_64k()
    length = 40
    rule = _as_f32_params(rule)
    q = torch.from_numpy(w.charges.astype(np.float64)).cuda()
    s = torch.from_numpy(w.seeds.astype(np.float64)).cuda()
    _fast_vs_exact(gpu, monkeypatch, q, s, length, rule, f"lines_64k {rules.NAMES[rule.rule]}")


def _fast_vs_exact(gpu, monkeypatch, q, s, length, rule, label):
    fast, fs = gpu.trace64(q, s, length, rule, with_steps=True)
    monkeypatch.setenv("EMTRACE_FP64_EXACT", "1")
    exact, es = gpu.trace64(q, s, length, rule, with_steps=True)
    monkeypatch.delenv("EMTRACE_FP64_EXACT")
    fast, exact = fast.cpu().numpy(), exact.cpu().numpy()
    same = (fast == exact) | (np.isnan(fast) & np.isnan(exact))
    print(f"[fp64 fast vs exact {label}] identical {same.mean():.8f}", flush=True)
    assert np.isfinite(exact).all()
    assert same.all() and torch.equal(fs, es)


@pytest.mark.parametrize("lanes", ["4", "8"])
def test_fp64_lane_groups_equal_compiler_expansions_default_problem(gpu, monkeypatch, lanes):
    """The reference application's default problem (1,024 charges, 16,384
    lines, 2,500 rows, curvature step) on doubles: four or eight lanes per
    line, one staged tile, each lane adding one component of its group's terms;
    bit for bit equal to the compiler's square roots and divisions with every
    lane adding all terms (4.2e10 terms)."""
    monkeypatch.setenv("EMTRACE_FP64_LANES", lanes)
    w = legacy.preset("default")
    q = torch.from_numpy(w.charges.astype(np.float64)).cuda()
    s = torch.from_numpy(w.seeds.astype(np.float64)).cuda()
    _fast_vs_exact(gpu, monkeypatch, q, s, w.len, _as_f32_params(rules.ref(1.0)),
                   f"default problem, {lanes} lanes")


@pytest.mark.parametrize("rule", [rules.plain(10.0), rules.angle()], ids=["plain", "angle"])
def test_fp64_four_lane_groups_staged_tiles_equal_compiler_expansions(gpu, monkeypatch, rule):
    """All 16,384 lines_64k charges (16 staged tiles per step) and every fourth
    of its lines (16,384: four lanes per line), 60 rows."""
    w = gen.lines_64k()
    q = torch.from_numpy(w.charges.astype(np.float64)).cuda()
    s = torch.from_numpy(np.ascontiguousarray(w.seeds[::4]).astype(np.float64)).cuda()
    _fast_vs_exact(gpu, monkeypatch, q, s, 60, _as_f32_params(rule), "lines_64k / 4")


def test_fp64_host_and_pitched_device_forms(gpu, orc):
    rule = _as_f32_params(rules.angle())
    w = gen.plain_config(200, 300, 5.0, 80, seed=4)
    ch, sd = w.charges.astype(np.float64), w.seeds.astype(np.float64)
    rc, want = orc.calc_lines(ch, sd, w.len, rule, dtype=np.float64)
    rows, perf, steps = gpu.trace_host64(ch, sd, w.len, rule)
    assert np.array_equal(rows, want)
    assert perf["progress"] == 1.0 and perf["kernel_s"] > 0 and perf["interactions"] == 200 * 300 * 79
    # device form with pitch > n: the padding columns are left alone
    import ctypes
    from emtrace import api
    from emtrace._lib import call
    pitch = 320
    buf = torch.full((3, w.len, pitch), 7.0, dtype=torch.float64, device="cuda")
    buf[:, 0, :w.n] = torch.from_numpy(sd).cuda().t()
    st = api.make_step(rule)
    q = torch.from_numpy(ch).cuda()
    call("em_trace_dev_f64", q.data_ptr(), w.p, buf[0].data_ptr(), buf[1].data_ptr(),
         buf[2].data_ptr(), w.n, w.len, pitch, ctypes.byref(st), None, api._stream_handle())
    torch.cuda.synchronize()
    got = buf.cpu().numpy()
    assert np.array_equal(got[:, :, :w.n], want)
    assert (got[:, 1:, w.n:] == 7.0).all()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "ref_f64_*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_fp64_vs_reference_double_golden(gpu, orc, path):
    g = np.load(path)
    res = float(g["resolution"])
    rule = rules.ref(res) if int(g["use_curvature"]) else rules.plain(res)
    length = g["rows"].shape[1]
    x = np.ascontiguousarray(np.repeat(g["seeds"][:, 0][None], length, 0))
    y = np.ascontiguousarray(np.repeat(g["seeds"][:, 1][None], length, 0))
    z = np.ascontiguousarray(np.repeat(g["seeds"][:, 2][None], length, 0))
    n = g["seeds"].shape[0]
    rc, perf = gpu.calc_lines64(g["charges"], x, y, z, n, res, bool(int(g["use_curvature"])))
    assert rc == 0 and perf["progress"] == 1.0
    got = np.stack([x, y, z])
    rc, want = orc.calc_lines(g["charges"], g["seeds"], length, rule, dtype=np.float64)
    assert np.array_equal(got, want)
    assert np.abs(got - g["rows"]).max() < 1e-9


@pytest.mark.parametrize("rule", [rules.ref(1.0), rules.plain(1.0)], ids=["ref", "plain"])
def test_fp64_host_path_row_chunks_match_device_path(gpu, rule):
    """The reference's default problem (1,024 charges, 16,384 lines, 2,500
    rows) in FP64 through the host-array entry point: 983 MB of output rows,
    traced in eight row chunks whose copies back overlap the next chunk's
    kernel.  Each chunk continues from the previous chunk's last row and the
    seed stays row 0, so the rows and step counts are bit-identical to one
    device-side launch; a second call reuses the cached device buffers."""
    w = legacy.preset("default")
    ch, sd = w.charges.astype(np.float64), w.seeds.astype(np.float64)
    rows_d, steps_d = gpu.trace64(torch.from_numpy(ch).cuda(), torch.from_numpy(sd).cuda(), w.len,
                                  rule, with_steps=True)
    rows_d, steps_d = rows_d.cpu().numpy(), steps_d.cpu().numpy()
    for call_no in range(2):
        rows_h, perf, steps_h = gpu.trace_host64(ch, sd, w.len, rule)
        print(f"[fp64 chunks {rule.rule}] call {call_no}: {perf['time_s'] * 1e3:.1f} ms, kernel "
              f"{perf['kernel_s'] * 1e3:.1f} ms, d2h {perf['d2h_s'] * 1e3:.1f} ms", flush=True)
        assert perf["progress"] == 1.0 and perf["kernel_s"] > 0
        assert np.array_equal(steps_h.astype(np.int64), steps_d.astype(np.int64))
        assert np.array_equal(rows_h, rows_d)
        del rows_h


def test_fp64_dropin_status_codes(gpu):
    ch = np.zeros((4, 4))
    x = np.zeros(64 * 10)
    y, z = x.copy(), x.copy()
    assert gpu.calc_lines64(ch, x, y, z, 0, 1.0, True)[0] == 1
    assert gpu.calc_lines64(ch, x, y, z, 64, 0.0, True)[0] == 2
    x1, y1, z1 = np.zeros(64), np.zeros(64), np.zeros(64)
    assert gpu.calc_lines64(ch, x1, y1, z1, 64, 1.0, True)[0] == 3


def test_fp32_kernel_vs_fp64_kernel_full_lines_8k(gpu, orc):
    """All of lines_8k on both kernels: the FP64 kernel stands in for the oracle at
    a size the CPU restatement would take minutes on.  FP32 rows drift from FP64
    over 2,500 steps on any hardware; the GPU's FP32 rows must drift no more
    than the CPU FP32 restatement's (sampled lines, prefixes before the first
    approach within 0.05 of a charge)."""
    from helpers import line_errors, prefix_lengths
    w = gen.lines_8k()
    rule = rules.ref(10.0)
    q32 = torch.from_numpy(w.charges).cuda()
    r32 = gpu.trace(q32, torch.from_numpy(w.seeds).cuda(), w.len, rule).cpu().numpy()
    r64 = gpu.trace64(q32.double(), torch.from_numpy(w.seeds).cuda().double(), w.len,
                      _as_f32_params(rule)).cpu().numpy()
    assert np.isfinite(r64).all()
    idx = np.arange(0, w.n, 16)
    rc, c32 = orc.calc_lines(w.charges, w.seeds[idx], w.len, rule)
    assert rc == 0
    ref64 = r64[:, :, idx]
    first, _ = prefix_lengths(ref64.astype(np.float32), w.charges)
    e_gpu = line_errors(r32[:, :, idx], ref64, first)
    e_cpu = line_errors(c32, ref64, first)
    q = lambda e, f: float(np.quantile(e, f))  # noqa: E731
    print(f"[lines_8k vs FP64 kernel, {idx.size} lines] gpu fp32 median {q(e_gpu, .5):.3e} p90 "
          f"{q(e_gpu, .9):.3e} max {e_gpu.max():.3e}; cpu fp32 median {q(e_cpu, .5):.3e} p90 "
          f"{q(e_cpu, .9):.3e} max {e_cpu.max():.3e}", flush=True)
    assert q(e_gpu, 0.5) <= 2 * q(e_cpu, 0.5) + 1e-6
    assert q(e_gpu, 0.9) <= 2 * q(e_cpu, 0.9) + 1e-6


def test_cli_double_precision_cpu_equals_gpu(gpu, tmp_path):
    import subprocess
    cli = os.path.join(os.path.dirname(GOLDEN), "..", "electromag-with-cuda_amd", "bin", "electromag")
    r = subprocess.run([cli, "--bogo", "--cpu", "--cpuprecision=double", "--gpuprecision=double",
                        "--save", "--autoregress", "--nodisp"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "data.txt").read().splitlines()
    cpu = [s[6:] for s in lines if s.startswith(" CPUL ")]
    gpu_ = [s[6:] for s in lines if s.startswith(" GPUL ")]
    assert len(cpu) == 256 * 500 and cpu == gpu_
    assert open(tmp_path / "regression.txt").read() == ""
    w = legacy.preset("bogo")
    assert lines[0].startswith(" CPUL [0][0] x: %g" % float(w.seeds[0, 0]))


def test_full_lines_64k_fp32_vs_fp64_kernel(gpu, orc):
    """The headline workload's charges and all 65,536 lines, first 250 rows,
    angle rule (all steps), on both kernels.  The FP32 kernel's distance to the
    FP64 kernel's polylines is no larger than the CPU FP32 path's (256 sampled
    lines; prefixes before the first approach within 0.05 of a charge)."""
    from helpers import prefix_lengths
    w = gen.lines_64k()
    length = 250
    rule = rules.angle(stop=False)
    q = torch.from_numpy(w.charges).cuda()
    s = torch.from_numpy(w.seeds).cuda()
    r32 = gpu.trace(q, s, length, rule).cpu().numpy()
    r64 = gpu.trace64(q.double(), s.double(), length, _as_f32_params(rule)).cpu().numpy()
    assert np.isfinite(r32).all() and np.isfinite(r64).all()
    idx = np.arange(0, w.n, 256)
    rc, c32 = orc.calc_lines(w.charges, w.seeds[idx], length, rule)
    assert rc == 0
    ref = np.ascontiguousarray(r64[:, :, idx]).astype(np.float32)
    first, _ = prefix_lengths(ref, w.charges)
    e_gpu = orc.path_dist(np.ascontiguousarray(r32[:, :, idx]), ref, first).max(axis=0)
    e_cpu = orc.path_dist(c32, ref, first).max(axis=0)
    # all lines: distance of the FP32 kernel's points to the FP64 kernel's polylines
    # (pointwise distances measure the arc-length phase of the adaptive step instead)
    r64f = r64.astype(np.float32)
    first_all, _ = prefix_lengths(r64f, w.charges)
    allerr = orc.path_dist(r32, r64f, first_all).max(axis=0)
    q_ = lambda e, f: float(np.quantile(e, f))  # noqa: E731
    print(f"[lines_64k 250 rows vs FP64 kernel] sampled {idx.size}: gpu path median "
          f"{q_(e_gpu, .5):.2e} p90 {q_(e_gpu, .9):.2e}; cpu path median {q_(e_cpu, .5):.2e} "
          f"p90 {q_(e_cpu, .9):.2e}; all {w.n} lines: path median {q_(allerr, .5):.2e} "
          f"p99 {q_(allerr, .99):.2e}", flush=True)
    assert q_(e_gpu, 0.5) <= 2 * q_(e_cpu, 0.5) + 1e-6
    assert q_(e_gpu, 0.9) <= 2 * q_(e_cpu, 0.9) + 1e-6


def test_full_lines_64k_field_vs_fp64(gpu, orc):
    """Per-evaluation field with all 16,384 lines_64k charges vs the FP64 sum."""
    w = gen.lines_64k()
    rng = np.random.default_rng(7)
    pts = np.concatenate([rng.uniform(-55, 55, (3000, 3)),
                          w.seeds[rng.integers(0, w.n, 1000)].astype(np.float64)]).astype(np.float32)
    E = gpu.field(torch.from_numpy(w.charges).cuda(), torch.from_numpy(pts).cuda(), 1.0)
    E = E.cpu().numpy().astype(np.float64)
    E64 = orc.field64(w.charges.astype(np.float64), pts.astype(np.float64), 1.0)
    c64 = w.charges.astype(np.float64)
    mag = np.concatenate([
        (np.abs(c64[None, :, 3]) / ((pts[k:k + 250].astype(np.float64)[:, None, :] - c64[None, :, :3])
                                    ** 2).sum(axis=2)).sum(axis=1)
        for k in range(0, len(pts), 250)])
    n64 = np.linalg.norm(E64, axis=1)
    rel = np.linalg.norm(E - E64, axis=1) / n64
    ok = n64 >= 0.1 * mag
    print(f"[lines_64k field] {ok.sum()} well-conditioned of {len(pts)}: max rel {rel[ok].max():.3g}, "
          f"median {np.median(rel[ok]):.3g}; all max |dE|/sum|E_i| "
          f"{(np.linalg.norm(E - E64, axis=1) / mag).max():.3g}", flush=True)
    assert ok.sum() > 500
    assert rel[ok].max() <= 1e-5
    # where the sum cancels: one FP32 term q d rsq(r^2)^3 carries up to ~10 ulp
    # (the rsq error is tripled by the cube), so |dE| <= 2e-6 sum|E_i|
    assert (np.linalg.norm(E - E64, axis=1) / mag).max() <= 2e-6
