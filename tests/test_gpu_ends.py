"""The angle rule with line ends (stop on, no checkpoints, at least
TAIL_MIN charges).  Up to WIDE_MAX charges it takes the register-resident
trace: 1,024-lane workgroups that hold every charge record in registers and
trace up to 16 lines at a time, in three passes (lines from a counter until
it runs out, then every workgroup's lines settled for up to 128 more steps,
then the lines still moving dealt out by work left and traced to the end).
Above WIDE_MAX it takes the lane-group trace: groups that take a new line
whenever theirs ends, up to a cap of CAP steps (DRAIN for the last lines
handed out), then 256-lane workgroups (two per CU) that trace the lines
still moving, up to 8 at a time -- or, when more lines than that
(TAIL_SLOTS on a 256-CU MI355X) reach the cap, the first phase's groups
continue them.  Both fill the rows after each line's end with its end point.
With fewer charges the same calls take the one-kernel path.

Parity: against the CPU path (identical step counts, points within 1e-4 on
well-conditioned lines, path distance to FP64 no worse than the CPU float32
path's), across the passes; bitwise against the one-kernel path where both
run the same arithmetic (below TAIL_MIN; the lane-group trace on rows up to
the drain cap), closely elsewhere.  The register-resident trace's rows do not
depend on the schedule: bitwise the same with fewer lines per workgroup and
fewer workgroups.  Also: determinism under dynamic line assignment, shapes at
the edges of the passes, many lines past the cap, pitch > n, graph replays,
and reuse of the work memory across sizes.
"""
import numpy as np
import pytest
import torch

from emtrace import gen, rules
from helpers import line_gates, rows64_to_soa, step_gate
from test_gpu_trace import dev

pytestmark = pytest.mark.gpu
CAP = 192  # lane-group trace: steps of the first phase (the second starts at step CAP + 1)
DRAIN = 128  # ... for the last lines handed out (every line, when no more lines than groups)
TAIL_MIN = 4096  # fewer charges: the one-kernel path
WIDE_MAX = 16384  # more charges: the lane-group trace
TAIL_SLOTS = 2 * 256 * 8  # lines the lane-group trace's second phase holds at once on 256 CUs


def escaping(p, n, half_side, length, seed):
    """All charges positive: lines escape and take every row (except the few
    that a neighbouring charge turns back next to their seed)."""
    w = gen.plain_config(p, n, half_side, length, seed=seed)
    ch = w.charges.copy()
    ch[:, 3] = np.abs(ch[:, 3])
    return ch, gen.make_seeds(ch, n, seed=seed)


def check_lines(orc, g_rows, c_rows, charges, seeds, length, tag, rule=None, min_conditioned=0.15):
    """The per-line gate of the one-kernel path's tests (tests/helpers.py,
    line_gates) against the FP64 trajectory.  Escaping lines travel far (|x| up
    to ~100, where one float32 ulp is 8e-6), so fewer lines are conditioned
    than near the charges; at least 15 % must be."""
    rule = rule or rules.angle()
    r64 = rows64_to_soa(orc.trace64(charges, seeds, length, rule))
    return line_gates(f"ends {tag}", g_rows, c_rows, r64, charges, seeds, rule,
                      min_conditioned=min_conditioned)


def ends_trace(gpu, charges, seeds, length, launch=None, pitch=None, rule=None):
    rows = gpu.seed_rows(dev(seeds), length, pitch)
    steps = torch.empty(seeds.shape[0], dtype=torch.int32, device="cuda")
    gpu.trace_into(dev(charges), rows, seeds.shape[0], rule or rules.angle(), launch, steps=steps)
    torch.cuda.synchronize()
    return rows.cpu().numpy(), steps.cpu().numpy().astype(np.int64)


def one_kernel_trace(gpu, charges, seeds, length, launch):
    """The same rule through the checkpointed forward (one kernel, no line refill)."""
    n = seeds.shape[0]
    rows = gpu.seed_rows(dev(seeds), length)
    every = 50
    ckpt = torch.empty(gpu.ckpt_size(n, length, every), dtype=torch.float32, device="cuda")
    steps = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu.trace_into(dev(charges), rows, n, rules.angle(), launch, ckpt=ckpt, ckpt_every=every,
                   steps=steps)
    torch.cuda.synchronize()
    return rows.cpu().numpy(), steps.cpu().numpy().astype(np.int64)


def lines_64k_subset(step=64):
    """lines_64k's 16,384 charges and every step-th of its seeds."""
    w = gen.lines_64k()
    return w.charges, w.seeds[::step]


def test_across_the_cap_vs_cpu_path(gpu, orc):
    """lines_64k charges, 1,024 of its lines, 600 rows: most lines end in the
    first phase, the escaping ones finish in the second."""
    charges, seeds = lines_64k_subset()
    w = gen.Workload(charges, seeds, 600)
    length = 600
    g_rows, g_steps = ends_trace(gpu, w.charges, seeds, length)
    rc, c_rows, c_steps = orc.calc_lines(w.charges, seeds, length, rules.angle(), with_steps=True)
    assert rc == 0
    past = (g_steps > CAP).sum()
    print(f"[ends] lines_64k subset/600: lines past the cap {past}/{len(g_steps)}, "
          f"identical steps {(g_steps == c_steps).sum()}", flush=True)
    assert past >= 16  # the second phase ran
    step_gate("ends lines_64k subset/600", w.charges, g_rows, g_steps, c_rows, c_steps)
    check_lines(orc, g_rows, c_rows, w.charges, seeds, length, "lines_64k subset/600")
    # rows after a line's end repeat its end point
    for l in range(len(g_steps)):
        e = g_steps[l]
        assert (g_rows[:, e + 1:, l] == g_rows[:, e:e + 1, l]).all()


@pytest.mark.parametrize("source", ["lines_8k", "lines_64k", "above_wide"])
def test_matches_one_kernel_path(gpu, source):
    """Against the one-kernel path (the checkpointed forward) with the same
    launch shape.  Below TAIL_MIN charges (lines_8k, 1,000) both calls take the
    one-kernel path and every row is bit-identical.  The register-resident
    trace (lines_64k's 16,384 charges) splits each charge sum differently, so
    its rows agree closely but not bitwise.  Above WIDE_MAX (lines_64k's
    charges and 600 more) the lane-group trace with fewer lines than lane
    groups gives every line the drain cap: rows 0..DRAIN and every line that
    ends before it are bit-identical, later rows agree closely."""
    if source == "lines_8k":
        w = gen.lines_8k()
        charges, seeds = w.charges, w.seeds[::8]
    elif source == "lines_64k":
        charges, seeds = lines_64k_subset()
    else:
        extra = gen.make_charges(WIDE_MAX + 600, 50.0, seed=3)[WIDE_MAX:]
        charges = np.concatenate([gen.lines_64k().charges, extra]).astype(np.float32)
        seeds = lines_64k_subset()[1]
    length = 400
    launch = gpu.make_launch(32, 512, 512)
    a_rows, a_steps = ends_trace(gpu, charges, seeds, length, launch)
    b_rows, b_steps = one_kernel_trace(gpu, charges, seeds, length, launch)
    if charges.shape[0] < TAIL_MIN:
        assert np.array_equal(a_steps, b_steps) and np.array_equal(a_rows, b_rows)
        return
    short = b_steps < DRAIN
    if charges.shape[0] > WIDE_MAX:
        assert np.array_equal(a_steps[short], b_steps[short])
        assert np.array_equal(a_rows[:, :DRAIN + 1], b_rows[:, :DRAIN + 1])
        assert np.array_equal(a_rows[:, :, short], b_rows[:, :, short])
    # rounding of the split sums: identical step counts on nearly every line,
    # and points that stay close along the way on most lines (a line that
    # passes near a charge amplifies the rounding; the CPU-path gates judge
    # every line)
    same = (a_steps == b_steps).mean()
    d = np.abs(a_rows[:, :64].astype(np.float64) - b_rows[:, :64]).max(axis=(0, 1))
    q50, q90 = np.quantile(d, 0.5), np.quantile(d, 0.9)
    print(f"[ends] vs one-kernel ({source}): identical steps {same:.4f}, per-line max |diff| "
          f"over rows 0..63: median {q50:.2e}, 90th percentile {q90:.2e}, max {d.max():.2e}",
          flush=True)
    assert same >= 0.99
    assert q50 <= 1e-4 and q90 <= 1e-3


def test_register_resident_rows_do_not_depend_on_the_schedule(gpu, monkeypatch):
    """The register-resident trace adds every line's charge sum in one fixed
    order whichever workgroup holds the line and whenever: fewer lines per
    workgroup (EMTRACE_WIDE_SLOTS) or fewer workgroups (EMTRACE_WIDE_GRID)
    give the same bits."""
    charges, seeds = lines_64k_subset(32)
    base = ends_trace(gpu, charges, seeds, 400)
    assert (base[1] > 192).any() and (base[1] < 100).any()
    for name, value in (("EMTRACE_WIDE_SLOTS", "6"), ("EMTRACE_WIDE_SLOTS", "32"),
                        ("EMTRACE_WIDE_GRID", "37")):
        monkeypatch.setenv(name, value)
        got = ends_trace(gpu, charges, seeds, 400)
        monkeypatch.delenv(name)
        print(f"[ends] schedule {name}={value}: rows and steps identical", flush=True)
        assert np.array_equal(got[0], base[0]) and np.array_equal(got[1], base[1])


def test_deterministic_under_line_refill(gpu):
    charges, seeds = lines_64k_subset(32)
    a = ends_trace(gpu, charges, seeds, 500)
    b = ends_trace(gpu, charges, seeds, 500)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# escaping lines with steps up to 0.05 stay within ~15 of the charges, where
# float32 resolves 1e-6 and the CPU float32 path follows FP64
SHORT_STEPS = rules.angle(hmax=0.05)


@pytest.mark.parametrize("p", [1, 40, 4095, 4096, 4097, 8193, 12000, 16384, 16385])
def test_charge_counts_vs_cpu_path(gpu, orc, p):
    """Charge counts on both sides of the one-kernel path's threshold (4,096),
    of the register-resident trace's record counts per lane (4, 8, 12, 16 per
    1,024 lanes: 4,097 and 8,193 are one record past a step) and of its
    limit (16,384; 16,385 takes the lane-group trace)."""
    length = 260
    # the box grows with p: about 40 charges per 1,000 unit^3, as at p = 40
    ch, seeds = escaping(p, 96, 5.0 * max(1.0, p / 40.0) ** (1.0 / 3.0), length, seed=31 + p)
    g_rows, g_steps = ends_trace(gpu, ch, seeds, length, rule=SHORT_STEPS)
    rc, c_rows, c_steps = orc.calc_lines(ch, seeds, length, SHORT_STEPS, with_steps=True)
    assert rc == 0
    step_gate(f"ends p={p}", ch, g_rows, g_steps, c_rows, c_steps)
    # dense positive charges can still turn a line back next to its seed
    assert (g_steps == length - 1).mean() >= 0.8
    if p == 1:  # radial straight lines: pointwise
        err = np.abs(g_rows.astype(np.float64) - c_rows).max()
        print(f"[ends p=1] max pointwise {err:.2e}", flush=True)
        assert err <= 1e-4
    else:
        check_lines(orc, g_rows, c_rows, ch, seeds, length, f"p={p}", SHORT_STEPS)


@pytest.mark.parametrize("p,lines", [(4200, 3000), (4200, 5000), (16500, 3000), (16500, 5000)],
                         ids=["wide-fewer-lines-than-slots", "wide-refills",
                              "lane-groups-second-phase-refills", "lane-groups-continuation"])
def test_many_lines_past_the_cap(gpu, orc, p, lines):
    """Escaping lines nearly all take every row.  4,200 charges (the
    register-resident trace): 3,000 lines fit the 4,096 slots of 256
    workgroups at once, so the counter runs out at the first refill; 5,000 do
    not, so workgroups take new lines as old ones end.  16,500 charges (the
    lane-group trace): 3,000 lines past the cap are refilled by the second
    phase's workgroups (8 slots each); 5,000 (more than TAIL_SLOTS) are
    continued by the first phase's groups instead."""
    length = CAP + 40
    ch, seeds = escaping(p, lines, 5.0 * (p / 40.0) ** (1.0 / 3.0), length, seed=5)
    g_rows, g_steps = ends_trace(gpu, ch, seeds, length, rule=SHORT_STEPS)
    rc, c_rows, c_steps = orc.calc_lines(ch, seeds, length, SHORT_STEPS, with_steps=True)
    assert rc == 0
    step_gate(f"ends {lines} escaping", ch, g_rows, g_steps, c_rows, c_steps)
    past = (g_steps > CAP).sum()
    assert past >= (2000 if lines == 3000 else TAIL_SLOTS + 1)
    check_lines(orc, g_rows, c_rows, ch, seeds, length, f"{lines} escaping", SHORT_STEPS)


@pytest.mark.parametrize("length", [2, CAP, CAP + 1, CAP + 2, CAP + 60])
def test_lengths_at_the_cap(gpu, orc, length):
    ch, seeds = escaping(4500, 130, 5.0 * (4500 / 40.0) ** (1.0 / 3.0), length, seed=7)
    ch[::3, 3] *= -1.0  # some sinks: a mix of ending and escaping lines
    g_rows, g_steps = ends_trace(gpu, ch, seeds, length)
    rc, c_rows, c_steps = orc.calc_lines(ch, seeds, length, rules.angle(), with_steps=True)
    assert rc == 0
    step_gate(f"ends len={length}", ch, g_rows, g_steps, c_rows, c_steps)
    if length > 2:
        check_lines(orc, g_rows, c_rows, ch, seeds, length, f"len={length}")
    else:
        assert np.isfinite(g_rows).all()


def test_no_charges(gpu, orc):
    """p = 0: E = 0, the direction is 0/0 and the line never reverses (NaN
    compares false), as on the CPU path; both phases run on NaN points."""
    length = CAP + 8
    w = gen.plain_config(4, 64, 5.0, length, seed=2)
    empty = np.zeros((0, 4), np.float32)
    g_rows, g_steps = ends_trace(gpu, empty, w.seeds, length)
    rc, c_rows, c_steps = orc.calc_lines(empty, w.seeds, length, rules.angle(), with_steps=True)
    assert rc == 0
    step_gate("ends no charges", empty, g_rows, g_steps, c_rows, c_steps)
    assert np.array_equal(g_rows[:, 0], c_rows[:, 0])
    assert np.isnan(g_rows[:, 1:]).all() and np.isnan(c_rows[:, 1:]).all()


def test_pitch_wider_than_n(gpu):
    """Rows with pitch > n: columns n..pitch-1 are not written."""
    w = gen.lines_64k()
    seeds = w.seeds[:200]
    length = 300
    ref_rows, ref_steps = ends_trace(gpu, w.charges, seeds, length)
    rows = gpu.seed_rows(dev(seeds), length, 237)
    rows[:, :, 200:] = 7.0
    steps = torch.empty(200, dtype=torch.int32, device="cuda")
    gpu.trace_into(dev(w.charges), rows, 200, rules.angle(), steps=steps)
    torch.cuda.synchronize()
    r = rows.cpu().numpy()
    assert (r[:, :, 200:] == 7.0).all()
    assert np.array_equal(r[:, :, :200], ref_rows)
    assert np.array_equal(steps.cpu().numpy().astype(np.int64), ref_steps)


def test_work_memory_reused_across_sizes(gpu):
    """A small call, a larger one (the work memory grows), the small one again."""
    w = gen.lines_64k()
    small = ends_trace(gpu, w.charges, w.seeds[:100], 250)
    big = ends_trace(gpu, w.charges, w.seeds[:4000], 250)
    again = ends_trace(gpu, w.charges, w.seeds[:100], 250)
    assert np.array_equal(small[0], again[0]) and np.array_equal(small[1], again[1])
    assert np.array_equal(big[0][:, :, :100], small[0]) and np.array_equal(big[1][:100], small[1])


def test_graph_replays_match_eager(gpu):
    """Captured into a graph (the counters are cleared by a kernel, not a
    memset), replayed three times with eager work in between: every replay
    equals the eager result."""
    charges_np, seeds = lines_64k_subset()
    length = 300
    n = seeds.shape[0]
    charges = dev(charges_np)
    eager_rows, eager_steps = ends_trace(gpu, charges_np, seeds, length)  # sizes the work memory
    w = gen.lines_8k()
    rows = gpu.seed_rows(dev(seeds), length)
    steps = torch.empty(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        gpu.trace_into(charges, rows, n, rules.angle(), steps=steps)  # warm-up on the side stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        gpu.trace_into(charges, rows, n, rules.angle(), steps=steps)
    for _ in range(3):
        rows[:, 1:].fill_(-1.0)
        steps.fill_(-1)
        junk = ends_trace(gpu, w.charges, w.seeds[:64], 100)  # eager work between replays
        assert np.isfinite(junk[0]).all()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(rows.cpu().numpy(), eager_rows)
        assert np.array_equal(steps.cpu().numpy().astype(np.int64), eager_steps)


def test_graph_replay_after_work_memory_grows(gpu):
    """A graph captured at one size still replays correctly after an eager call
    has grown the work memory past anything used before (the outgrown buffer
    the graph holds is retired, not freed)."""
    charges_np, seeds = lines_64k_subset(128)
    length = 250
    n = seeds.shape[0]
    charges = dev(charges_np)
    eager_rows, eager_steps = ends_trace(gpu, charges_np, seeds, length)
    w = gen.lines_8k()
    rows = gpu.seed_rows(dev(seeds), length)
    steps = torch.empty(n, dtype=torch.int32, device="cuda")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        gpu.trace_into(charges, rows, n, rules.angle(), steps=steps)
    # 250,000 lines: more work memory than any other test asks for
    big_seeds = np.tile(w.seeds, (31, 1))[:250000]
    big = gpu.seed_rows(dev(big_seeds), 3)
    gpu.trace_into(charges, big, big_seeds.shape[0], rules.angle())
    torch.cuda.synchronize()
    assert bool(torch.isfinite(big).all())
    for _ in range(2):
        rows[:, 1:].fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(rows.cpu().numpy(), eager_rows)
        assert np.array_equal(steps.cpu().numpy().astype(np.int64), eager_steps)
