"""The two-phase trace of the angle rule with line ends (stop on, no
checkpoints): lane groups that take a new line whenever theirs ends, up to a
cap of CAP steps; then one 256-lane workgroup per line still moving; then the
rows after each line's end are filled with its end point.

Parity: against the CPU path (identical step counts, points within 1e-4 on
well-conditioned lines, path distance to FP64 no worse than the CPU float32
path's), across the cap; bitwise against the one-kernel path (the
checkpointed forward, same launch shape) on rows up to the cap, where both
run the same arithmetic in the same order.  Also: determinism under dynamic
line assignment, shapes at the edges of both phases, more lines set aside than
the second phase holds at once, pitch > n, graph replays, and reuse of the
work memory across sizes.
"""
import numpy as np
import pytest
import torch

from emtrace import gen, rules
from helpers import rows64_to_soa
from test_gpu_trace import check_against_cpu, dev

pytestmark = pytest.mark.gpu
CAP = 192  # steps of the first phase (the second phase starts at step CAP + 1)


def escaping(p, n, half_side, length, seed):
    """All charges positive: lines escape and take every row (except the few
    that a neighbouring charge turns back next to their seed)."""
    w = gen.plain_config(p, n, half_side, length, seed=seed)
    ch = w.charges.copy()
    ch[:, 3] = np.abs(ch[:, 3])
    return ch, gen.make_seeds(ch, n, seed=seed)


def check_lines(orc, g_rows, c_rows, charges, seeds, length, tag, rule=None):
    """The angle rule's bar (as for the one-kernel path): path distance to the
    CPU path on lines whose CPU float32 path follows FP64, and distance to the
    FP64 path no worse than the CPU float32 path's."""
    rule = rule or rules.angle()
    r64 = rows64_to_soa(orc.trace64(charges, seeds, length, rule))
    check_against_cpu(f"ends {tag}", g_rows, c_rows, r64, charges, rule)


def same_steps(g_steps, c_steps):
    """Step counts equal to the CPU path's on all but rounding-decided lines
    (a direction reversal whose dot product is within rounding of zero)."""
    same = g_steps == c_steps.astype(np.int64)
    return same.mean() >= 0.97


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


def test_lines_8k_across_the_cap_vs_cpu_path(gpu, orc):
    """lines_8k charges, 1,024 of its lines, 600 rows: most lines end in the
    first phase, the escaping ones finish in the second."""
    w = gen.lines_8k()
    seeds = w.seeds[::8]
    length = 600
    g_rows, g_steps = ends_trace(gpu, w.charges, seeds, length)
    rc, c_rows, c_steps = orc.calc_lines(w.charges, seeds, length, rules.angle(), with_steps=True)
    assert rc == 0
    past = (g_steps > CAP).sum()
    print(f"[ends] lines_8k/600: lines past the cap {past}/{len(g_steps)}, "
          f"identical steps {(g_steps == c_steps).sum()}", flush=True)
    assert past >= 16  # the second phase ran
    assert (g_steps == c_steps.astype(np.int64)).mean() >= 0.99
    r64 = rows64_to_soa(orc.trace64(w.charges, seeds, length, rules.angle()))
    check_against_cpu("ends lines_8k/600", g_rows, c_rows, r64, w.charges, rules.angle())
    # rows after a line's end repeat its end point
    for l in range(len(g_steps)):
        e = g_steps[l]
        assert (g_rows[:, e + 1:, l] == g_rows[:, e:e + 1, l]).all()


def test_matches_one_kernel_path_bitwise_up_to_the_cap(gpu):
    """Same launch shape: rows 0..CAP and every step count below the cap are
    bit-identical to the one-kernel path; later rows agree closely."""
    w = gen.lines_8k()
    seeds = w.seeds[::8]
    length = 400
    launch = gpu.make_launch(32, 512, 512)
    a_rows, a_steps = ends_trace(gpu, w.charges, seeds, length, launch)
    b_rows, b_steps = one_kernel_trace(gpu, w.charges, seeds, length, launch)
    short = b_steps < CAP
    assert np.array_equal(a_steps[short], b_steps[short])
    assert np.array_equal(a_rows[:, :CAP + 1], b_rows[:, :CAP + 1])
    assert np.array_equal(a_rows[:, :, short], b_rows[:, :, short])
    long_ = ~short
    d = np.abs(a_rows[:, :, long_].astype(np.float64) - b_rows[:, :, long_]).max()
    print(f"[ends] vs one-kernel: {long_.sum()} lines past the cap, max |diff| {d:.2e}", flush=True)
    assert long_.sum() >= 16


def test_deterministic_under_line_refill(gpu):
    w = gen.lines_8k()
    seeds = w.seeds[::4]
    a = ends_trace(gpu, w.charges, seeds, 500)
    b = ends_trace(gpu, w.charges, seeds, 500)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# escaping lines with steps up to 0.05 stay within ~15 of the charges, where
# float32 resolves 1e-6 and the CPU float32 path follows FP64
SHORT_STEPS = rules.angle(hmax=0.05)


@pytest.mark.parametrize("p", [1, 40, 4095, 4096, 4097, 12000])
def test_charge_counts_vs_cpu_path(gpu, orc, p):
    """Charge counts around the second phase's padding granule (4,096) and
    beyond one LDS tile (12,000: the first phase re-stages the charges)."""
    length = 260
    # the box grows with p: about 40 charges per 1,000 unit^3, as at p = 40
    ch, seeds = escaping(p, 96, 5.0 * max(1.0, p / 40.0) ** (1.0 / 3.0), length, seed=31 + p)
    g_rows, g_steps = ends_trace(gpu, ch, seeds, length, rule=SHORT_STEPS)
    rc, c_rows, c_steps = orc.calc_lines(ch, seeds, length, SHORT_STEPS, with_steps=True)
    assert rc == 0
    assert same_steps(g_steps, c_steps)
    # dense positive charges can still turn a line back next to its seed
    assert (g_steps == length - 1).mean() >= 0.8
    if p == 1:  # radial straight lines: pointwise
        err = np.abs(g_rows.astype(np.float64) - c_rows).max()
        print(f"[ends p=1] max pointwise {err:.2e}", flush=True)
        assert err <= 1e-4
    else:
        check_lines(orc, g_rows, c_rows, ch, seeds, length, f"p={p}", SHORT_STEPS)


def test_more_lines_set_aside_than_second_phase_slots(gpu, orc):
    """3,000 escaping lines all reach the cap: the second phase (a few
    workgroups per CU) takes them in several rounds."""
    length = CAP + 40
    ch, seeds = escaping(64, 3000, 5.0, length, seed=5)
    g_rows, g_steps = ends_trace(gpu, ch, seeds, length, rule=SHORT_STEPS)
    rc, c_rows, c_steps = orc.calc_lines(ch, seeds, length, SHORT_STEPS, with_steps=True)
    assert rc == 0
    assert same_steps(g_steps, c_steps)
    assert (g_steps == length - 1).sum() >= 2000  # more than the second phase holds at once
    check_lines(orc, g_rows, c_rows, ch, seeds, length, "3000 escaping", SHORT_STEPS)


@pytest.mark.parametrize("length", [2, CAP, CAP + 1, CAP + 2])
def test_lengths_at_the_cap(gpu, orc, length):
    ch, seeds = escaping(50, 130, 5.0, length, seed=7)
    ch[::3, 3] *= -1.0  # some sinks: a mix of ending and escaping lines
    g_rows, g_steps = ends_trace(gpu, ch, seeds, length)
    rc, c_rows, c_steps = orc.calc_lines(ch, seeds, length, rules.angle(), with_steps=True)
    assert rc == 0
    assert same_steps(g_steps, c_steps)
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
    assert same_steps(g_steps, c_steps)
    assert np.array_equal(g_rows[:, 0], c_rows[:, 0])
    assert np.isnan(g_rows[:, 1:]).all() and np.isnan(c_rows[:, 1:]).all()


def test_pitch_wider_than_n(gpu):
    """Rows with pitch > n: columns n..pitch-1 are not written."""
    w = gen.lines_8k()
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
    w = gen.lines_8k()
    small = ends_trace(gpu, w.charges, w.seeds[:100], 250)
    big = ends_trace(gpu, w.charges, w.seeds[:4000], 250)
    again = ends_trace(gpu, w.charges, w.seeds[:100], 250)
    assert np.array_equal(small[0], again[0]) and np.array_equal(small[1], again[1])
    assert np.array_equal(big[0][:, :, :100], small[0]) and np.array_equal(big[1][:100], small[1])


def test_graph_replays_match_eager(gpu):
    """Captured into a graph (the counters are cleared by a kernel, not a
    memset), replayed three times with eager work in between: every replay
    equals the eager result."""
    w = gen.lines_8k()
    seeds = w.seeds[::8]
    length = 300
    n = seeds.shape[0]
    charges = dev(w.charges)
    eager_rows, eager_steps = ends_trace(gpu, w.charges, seeds, length)  # sizes the work memory
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
    w = gen.lines_8k()
    seeds = w.seeds[::16]
    length = 250
    n = seeds.shape[0]
    charges = dev(w.charges)
    eager_rows, eager_steps = ends_trace(gpu, w.charges, seeds, length)
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
