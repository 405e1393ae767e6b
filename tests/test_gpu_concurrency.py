"""Calls from several host threads at once (the reference's own threading
model: a synchronous calc-lines call per thread, and a monitor polling the
progress field).  Host-array calls share per-device buffers behind a lock;
device-path calls on distinct streams share the line-ends work memory and the
adjoint's work lists, ordered by events.  Every concurrent result must equal
the same call made alone, bit for bit."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from emtrace import gen, rules
from test_gpu_trace import dev

pytestmark = pytest.mark.gpu


def _drop_in(gpu, w):
    rows = np.zeros((3, w.len, w.n), np.float32)
    rows[:, 0] = w.seeds.T
    rc, perf = gpu.calc_lines(w.charges, rows[0], rows[1], rows[2], w.n, 10.0, True)
    assert rc == 0 and perf["progress"] == 1.0
    return rows


def _device(gpu, w, rule, stream):
    with torch.cuda.stream(stream):
        rows = gpu.seed_rows(dev(w.seeds), w.len)
        steps = torch.empty(w.n, dtype=torch.int32, device="cuda")
        gpu.trace_into(dev(w.charges), rows, w.n, rule, None, steps=steps)
        stream.synchronize()
        return rows.cpu().numpy(), steps.cpu().numpy()


def test_drop_in_calls_from_threads(gpu):
    ws = [gen.plain_config(300 + 100 * i, 640, 10.0, 120, seed=40 + i) for i in range(6)]
    alone = [_drop_in(gpu, w) for w in ws]
    with ThreadPoolExecutor(max_workers=6) as ex:
        together = list(ex.map(lambda w: _drop_in(gpu, w), ws))
    for a, b in zip(alone, together):
        assert np.array_equal(a, b)


def test_device_calls_on_streams_from_threads(gpu):
    """Line ends on the register-resident trace (4,500 charges: shared work memory) and
    all-steps traces, each thread on its own stream."""
    ws = [gen.plain_config(4500, 600, 12.0, 260, seed=50 + i) for i in range(3)]
    ws += [gen.plain_config(800, 900, 10.0, 150, seed=60 + i) for i in range(3)]
    rls = [rules.angle()] * 3 + [rules.angle(stop=False)] * 3
    streams = [torch.cuda.Stream() for _ in ws]
    alone = [_device(gpu, w, r, s) for w, r, s in zip(ws, rls, streams)]
    with ThreadPoolExecutor(max_workers=len(ws)) as ex:
        together = list(ex.map(lambda a: _device(gpu, *a), zip(ws, rls, streams)))
    for (ra, sa), (rb, sb) in zip(alone, together):
        assert np.array_equal(sa, sb) and np.array_equal(ra, rb)


def test_adjoint_calls_on_streams_from_threads(gpu):
    """The single-configuration adjoint from several threads, two problems
    alternating (so calls rebuild and reuse the shared work lists)."""
    ws = [gen.plain_config(300, 200, 6.0, 90, seed=5), gen.plain_config(120, 96, 4.0, 50, seed=6)]
    rule = rules.angle()

    def run(i):
        w = ws[i % 2]
        G = np.random.default_rng(i % 2).normal(size=(3, w.len, w.n)).astype(np.float32)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            c = dev(w.charges)
            rows = gpu.seed_rows(dev(w.seeds), w.len)
            ck = torch.empty(gpu.ckpt_size(w.n, w.len, 7), dtype=torch.float32, device="cuda")
            gpu.trace_into(c, rows, w.n, rule, None, ck, 7)
            gc, gs = gpu.adjoint(c, rows, ck, 7, dev(G), rule)
            s.synchronize()
            return gc.cpu().numpy(), gs.cpu().numpy()

    alone = [run(i) for i in range(2)]
    with ThreadPoolExecutor(max_workers=6) as ex:
        together = list(ex.map(run, range(12)))
    for i, (gc, gs) in enumerate(together):
        assert np.array_equal(gc, alone[i % 2][0]) and np.array_equal(gs, alone[i % 2][1]), i
