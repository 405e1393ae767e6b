"""Batched form: many configurations with uneven charge and line counts in one
persistent launch, optionally replayed from a hipGraph.  Each configuration's
rows must equal, bit for bit, what the single-configuration kernel writes with
the same launch shape (the work split per line is identical)."""
import numpy as np
import pytest
import torch

from emtrace import gen, rules

pytestmark = pytest.mark.gpu


def _configs(count=12, seed=3, length=120):
    return gen.batched_configs(count, (64, 3000), (40, 700), length, seed)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_equals_single_configuration(gpu, use_graph):
    ws = _configs()
    layout, charges, lines = gpu.pack_batch(ws)
    plan = gpu.BatchPlan(layout, rules.angle(), None, use_graph=use_graph)
    info = plan.info()
    plan.run(charges, lines)
    torch.cuda.synchronize()
    launch = gpu.make_launch(info["lanes_per_line"], info["tile"], info["block"])
    for i, w in enumerate(ws):
        single = gpu.trace(torch.from_numpy(w.charges).cuda(), torch.from_numpy(w.seeds).cuda(),
                           w.len, rules.angle(), launch)
        got = layout.rows_view(lines, i)
        assert torch.equal(got, single), i
    assert abs(info["interactions"] - layout.interactions()) < 1


def test_batched_default_shape_with_a_large_configuration(gpu):
    """A configuration of 10,000 charges next to small ones: the default tile
    (sized for four workgroups per CU beside the kernels' static LDS) re-stages
    its charges every step; rows bitwise equal to the single-configuration
    kernel with the plan's shape."""
    ws = _configs(count=3, seed=12, length=30) + [gen.plain_config(10000, 100, 10.0, 30, seed=13)]
    layout, charges, lines = gpu.pack_batch(ws)
    plan = gpu.BatchPlan(layout, rules.angle(stop=False), None, use_graph=False)
    info = plan.info()
    assert info["tile"] < 10000 and info["tile"] * 16 + 512 <= 160 * 1024
    plan.run(charges, lines)
    torch.cuda.synchronize()
    launch = gpu.make_launch(info["lanes_per_line"], info["tile"], info["block"])
    for i, w in enumerate(ws):
        single = gpu.trace(torch.from_numpy(w.charges).cuda(), torch.from_numpy(w.seeds).cuda(),
                           w.len, rules.angle(stop=False), launch)
        assert torch.equal(layout.rows_view(lines, i), single), i
    plan.close()


def test_graph_replays_match_eager(gpu):
    ws = _configs(count=9, seed=4, length=80)
    layout, charges, lines = gpu.pack_batch(ws)
    eager = gpu.BatchPlan(layout, rules.angle(), None, use_graph=False)
    eager.run(charges, lines)
    torch.cuda.synchronize()
    want = lines.clone()
    graphed = gpu.BatchPlan(layout, rules.angle(), None, use_graph=True)
    for rep in range(4):
        lines[:, :] = 0.0
        for i, w in enumerate(ws):
            layout.rows_view(lines, i)[:, 0, :] = torch.from_numpy(w.seeds.T.copy()).cuda()
        torch.cuda.synchronize()
        graphed.run(charges, lines)
        torch.cuda.synchronize()
        x = lines.sum().item()  # eager work between replays
        assert np.isfinite(x)
        assert torch.equal(lines, want), rep


def test_batched_adjoint_equals_single(gpu):
    ws = _configs(count=5, seed=6, length=60)
    layout, charges, lines = gpu.pack_batch(ws)
    plan = gpu.BatchPlan(layout, rules.angle(), None, use_graph=False)
    info = plan.info()
    every = 9
    ck = torch.empty(plan.ckpt_size(every), dtype=torch.float32, device="cuda")
    plan.run_ckpt(charges, lines, ck, every)
    G = torch.randn_like(lines)
    grad = torch.empty_like(charges)
    plan.adjoint(charges, lines, ck, every, G, grad)
    torch.cuda.synchronize()
    launch = gpu.make_launch(info["lanes_per_line"], info["tile"], info["block"])
    for i, w in enumerate(ws):
        c = torch.from_numpy(w.charges).cuda()
        rows = gpu.seed_rows(torch.from_numpy(w.seeds).cuda(), w.len)
        ck1 = torch.empty(gpu.ckpt_size(w.n, w.len, every), dtype=torch.float32, device="cuda")
        gpu.trace_into(c, rows, w.n, rules.angle(), launch, ck1, every)
        g_rows = layout.rows_view(G, i).contiguous()
        gc, _ = gpu.adjoint(c, rows, ck1, every, g_rows, rules.angle(), launch)
        o = int(layout.charge_off[i])
        got = grad[o:o + w.p]
        assert torch.equal(rows, layout.rows_view(lines, i)), i
        # the charge pass chunks its sources per launch, so sums may differ in the last bits
        assert torch.allclose(got, gc, rtol=1e-4, atol=1e-6 * gc.abs().max().item()), i


def test_full_size_lines_64k_properties(gpu):
    """At the headline size: finite output, deterministic, every row written."""
    w = gen.lines_64k()
    c = torch.from_numpy(w.charges).cuda()
    s = torch.from_numpy(w.seeds).cuda()
    rows = gpu.seed_rows(s, w.len)
    rows[:, 1:, :] = float("nan")
    gpu.trace_into(c, rows, w.n, rules.angle())
    torch.cuda.synchronize()
    assert bool(torch.isfinite(rows).all())
    a = rows[:, -1, :].clone()
    gpu.trace_into(c, rows, w.n, rules.angle())
    torch.cuda.synchronize()
    assert torch.equal(rows[:, -1, :], a)


def test_full_size_lines_64k_invariances(gpu):
    """At the headline size with the shipped launch shape (4 lanes, tile 4,096,
    512 threads, unroll 2), all 2,499 steps: doubling every charge doubles
    every partial sum exactly, so the rows are bitwise unchanged; unroll 8
    consumes the records in the same order, so its rows are bitwise equal."""
    w = gen.lines_64k()
    c = torch.from_numpy(w.charges).cuda()
    s = torch.from_numpy(w.seeds).cuda()
    rule = rules.angle(stop=False)
    shipped = gpu.make_launch(4, 4096, 512, 2)
    a = gpu.trace(c, s, w.len, rule, shipped)
    c2 = c.clone()
    c2[:, 3] *= 2.0
    b = gpu.trace(c2, s, w.len, rule, shipped)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    del b
    u8 = gpu.trace(c, s, w.len, rule, gpu.make_launch(4, 4096, 512, 8))
    torch.cuda.synchronize()
    assert torch.equal(a, u8)


def test_full_size_batch_512(gpu):
    """The 512 configurations (1,000 rows) in one graph-replayed persistent
    launch with the library's default shape: finite, every line all 999 steps,
    and the configurations with the most and the fewest charges and lines
    bitwise equal to single-configuration traces with the plan's shape."""
    ws = gen.batch_512()
    layout, charges, lines = gpu.pack_batch(ws)
    rule = rules.angle(stop=False)
    plan = gpu.BatchPlan(layout, rule, None, use_graph=True)
    info = plan.info()
    steps = torch.empty(int(layout.n.sum()), dtype=torch.int32, device="cuda")
    plan.run(charges, lines, steps)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(lines).all())
    assert bool((steps == 999).all())
    launch = gpu.make_launch(info["lanes_per_line"], info["tile"], info["block"], info["unroll"])
    picks = {int(np.argmax(layout.p)), int(np.argmin(layout.p)), int(np.argmax(layout.n)),
             int(np.argmin(layout.n))}
    for i in sorted(picks):
        w = ws[i]
        single = gpu.trace(torch.from_numpy(w.charges).cuda(), torch.from_numpy(w.seeds).cuda(),
                           w.len, rule, launch)
        assert torch.equal(layout.rows_view(lines, i), single), i
    plan.close()


def test_autotuner_picks_a_valid_shape(gpu, tmp_path, monkeypatch):
    """The tuned shape is one of the candidates, is cached, and gives the same
    lines as an explicit launch with that shape."""
    from emtrace import autotune
    monkeypatch.setenv("EMTRACE_TUNE_CACHE", str(tmp_path / "tune.json"))
    w = gen.plain_config(1500, 3000, 10.0, 64, seed=9)
    c = torch.from_numpy(w.charges).cuda()
    s = torch.from_numpy(w.seeds).cuda()
    launch, report = autotune.tune_plain(c, s, rules.angle(), rows=16)
    shape = (launch.lanes_per_line, launch.block, launch.tile, launch.unroll)
    assert shape in [r[:4] for r in report] and all(np.isfinite(r[4]) for r in report)
    assert {r[3] for r in report} == {0, 2, 4, 8}  # default, then each unroll
    again, rep2 = autotune.tune_plain(c, s, rules.angle(), rows=16)
    assert (again.lanes_per_line, again.block, again.tile, again.unroll) == shape
    assert rep2[0][-1] == "cached"
    a = gpu.trace(c, s, w.len, rules.angle(), launch)
    b = gpu.trace(c, s, w.len, rules.angle(), gpu.make_launch(shape[0], shape[2], shape[1], shape[3]))
    assert torch.equal(a, b)
    blaunch, brep = autotune.tune_batch(_configs(count=6, seed=2, length=40), rules.angle(), rows=8)
    assert blaunch.lanes_per_line >= 1 and any(np.isfinite(r[4]) for r in brep)


def test_captured_checkpointed_graph_survives_plain_runs(gpu):
    """run_ckpt + adjoint captured into a graph; plain (checkpoint-free) eager
    runs of the same plan between replays leave the captured descriptors
    alone, so every replay equals the eager forward + adjoint."""
    ws = _configs(count=6, seed=8, length=70)
    layout, charges, lines = gpu.pack_batch(ws)
    plan = gpu.BatchPlan(layout, rules.angle(), None, use_graph=False)
    every = 10
    ck = torch.empty(plan.ckpt_size(every), dtype=torch.float32, device="cuda")
    G = torch.randn_like(lines)
    grad = torch.empty_like(charges)
    seeded = lines.clone()
    plan.run_ckpt(charges, lines, ck, every)
    plan.adjoint(charges, lines, ck, every, G, grad)
    torch.cuda.synchronize()
    want_lines, want_grad = lines.clone(), grad.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.run_ckpt(charges, lines, ck, every)
        plan.adjoint(charges, lines, ck, every, G, grad)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run_ckpt(charges, lines, ck, every)
        plan.adjoint(charges, lines, ck, every, G, grad)
    for rep in range(3):
        plain = seeded.clone()
        plan.run(charges, plain)  # eager, no checkpoints
        torch.cuda.synchronize()
        lines.copy_(seeded)
        grad.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(lines, want_lines), rep
        assert torch.equal(grad, want_grad), rep
        assert torch.equal(plain, want_lines), rep
