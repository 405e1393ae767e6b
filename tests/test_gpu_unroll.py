"""The unroll U of a launch shape (charge records per pipelined group, 2, 4 or
8) changes only how far ahead each lane's LDS reads run: every lane still
consumes its records in the same order and folds its compensated block sums at
the same records, so every unroll gives bit-identical rows, step counts and
checkpoints.  Covered: all steps and line ends (one-kernel and two-phase
paths), the checkpointed forward, the batched plan, and the rejection of any
other unroll."""
import numpy as np
import pytest
import torch

from emtrace import gen, rules
from emtrace._lib import EmError
from test_gpu_trace import dev

pytestmark = pytest.mark.gpu
UNROLLS = (2, 4, 8)


def trace_all(gpu, charges, seeds, length, rule, launch, ckpt_every=0):
    n = seeds.shape[0]
    rows = gpu.seed_rows(dev(seeds), length)
    steps = torch.empty(n, dtype=torch.int32, device="cuda")
    ckpt = None
    if ckpt_every:
        ckpt = torch.empty(gpu.ckpt_size(n, length, ckpt_every), dtype=torch.float32, device="cuda")
    gpu.trace_into(dev(charges), rows, n, rule, launch, ckpt=ckpt, ckpt_every=ckpt_every,
                   steps=steps)
    torch.cuda.synchronize()
    return rows.cpu(), steps.cpu(), None if ckpt is None else ckpt.cpu()


def assert_same(a, b):
    for x, y in zip(a, b):
        assert (x is None) == (y is None)
        if x is not None:
            assert torch.equal(x, y)


@pytest.mark.parametrize("S,block", [(1, 256), (4, 512), (32, 512)])
def test_all_steps_identical_for_every_unroll(gpu, S, block):
    """1,500 charges (a partial last tile), angle and plain rules."""
    w = gen.plain_config(1500, 2000, 10.0, 48, seed=21)
    for rule in (rules.angle(stop=False), rules.plain()):
        base = trace_all(gpu, w.charges, w.seeds, w.len, rule, gpu.make_launch(S, 0, block, 8))
        for U in UNROLLS[:-1]:
            got = trace_all(gpu, w.charges, w.seeds, w.len, rule, gpu.make_launch(S, 0, block, U))
            assert_same(base, got)


@pytest.mark.parametrize("p", [900, 5000])
def test_line_ends_identical_for_every_unroll(gpu, p):
    """Below 4,096 charges the one-kernel path, above it the register-resident
    trace (which takes no launch shape: its rows are the same for every
    unroll by construction); hmax 0.05 so that lines end and some outlast its
    settling pass within 400 rows."""
    w = gen.plain_config(p, 1200, 10.0, 400, seed=22)
    rule = rules.angle(hmax=0.05)
    base = trace_all(gpu, w.charges, w.seeds, w.len, rule, gpu.make_launch(32, 0, 512, 8))
    ended = base[1] < w.len - 1
    assert ended.any() and (~ended).any()
    for U in UNROLLS[:-1]:
        got = trace_all(gpu, w.charges, w.seeds, w.len, rule, gpu.make_launch(32, 0, 512, U))
        assert_same(base, got)


def test_checkpointed_forward_identical_for_every_unroll(gpu):
    w = gen.plain_config(2100, 1000, 10.0, 120, seed=23)
    rule = rules.angle(stop=True)
    base = trace_all(gpu, w.charges, w.seeds, w.len, rule, gpu.make_launch(4, 0, 256, 8), 25)
    for U in UNROLLS[:-1]:
        got = trace_all(gpu, w.charges, w.seeds, w.len, rule, gpu.make_launch(4, 0, 256, U), 25)
        assert_same(base, got)


def test_batch_identical_for_every_unroll(gpu):
    ws = gen.batched_configs(8, (64, 3000), (40, 700), 60, 24)
    layout, charges, lines = gpu.pack_batch(ws)
    out = []
    for U in UNROLLS:
        plan = gpu.BatchPlan(layout, rules.angle(), gpu.make_launch(4, 0, 256, U), use_graph=False)
        assert plan.info()["unroll"] == U
        work = lines.clone()
        plan.run(charges, work)
        torch.cuda.synchronize()
        out.append(work.cpu())
        plan.close()
    assert torch.equal(out[0], out[1]) and torch.equal(out[0], out[2])


@pytest.mark.parametrize("U", [1, 3, 16, -2])
def test_other_unrolls_are_rejected(gpu, U):
    w = gen.plain_config(300, 100, 10.0, 8, seed=25)
    with pytest.raises(EmError) as e:
        trace_all(gpu, w.charges, w.seeds, w.len, rules.angle(), gpu.make_launch(4, 0, 256, U))
    assert "unroll" in str(e.value)
