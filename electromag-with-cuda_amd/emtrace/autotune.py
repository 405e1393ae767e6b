"""Launch-shape autotuner for the tracer.

A launch shape is (lanes per line S, threads per workgroup, charges per LDS
tile, unroll U = charge records per pipelined group).  The tuner times each
(S, block, tile) candidate on the caller's own inputs with a reduced row count
(the per-step cost is what differs between shapes) at the library's default
unroll, then the fastest of them at U = 8, 4 and 2, and keeps the fastest.
The unroll never changes the rows.  The lanes per line S and the tile do:
S changes how each charge sum is split over lanes, and the tile where the
compensated partial sums are folded when the charges are re-staged, so rows
from two shapes agree to float32 summation rounding, not bit for bit.  A
forward and its adjoint must therefore use the same shape (autograd keeps the
forward's).  Results are cached in a JSON file keyed by device,
problem size bucket and step rule (``$EMTRACE_TUNE_CACHE``, default
``~/.cache/emtrace/tune.json``); a small table of shapes measured on MI355X
(``SHIPPED``) answers when the cache has no entry.

Candidates are generated from the device's LDS budget: lanes per line S in
{1, 2, ..., 32} (single configuration: until the grid holds 8 waves per SIMD;
batch: 1-8), blocks of 256 and 512 threads, and tiles that hold all charges
when they fit, else the largest tile that leaves room for one, two or four
resident workgroups per CU (also eight for the batch's persistent grid, where
more resident workgroups keep the SIMDs busy while items of uneven size end).
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, List, Optional, Tuple

from . import api
from .rules import StepRule

LDS_BYTES = 160 * 1024
LDS_STATIC = 512  # static LDS of the trace kernels (256-272 B), kept free beside the tile
GROUP = 16  # charges per lane per pipelined pair: tiles are multiples of S * GROUP


def _cache_path() -> str:
    return os.environ.get("EMTRACE_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "emtrace", "tune.json"))


def _load() -> Dict[str, list]:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, value: list) -> None:
    data = _load()
    data[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def _device_key() -> str:
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = str(getattr(p, "gcnArchName", p.name)).split(":")[0]
    return f"{arch}:{p.multi_processor_count}"


# Shapes measured by full sweeps on MI355X (gfx950, 256 CUs), used when the
# local cache has no entry: the fastest shape of each sweep, ties (within
# 1%) going to fewer lanes per line.  [S, block, tile, U, ms at 48 rows]
SHIPPED = {
    # 16,384 charges, 65,536 lines, angle rule, all steps (2,500 rows: 831 ms;
    # unroll 4: 847 ms, 8: 917 ms; 256-thread blocks: 984 ms at best).  With
    # line ends this problem takes the register-resident trace, which uses no
    # launch shape (tune_plain does not time it).
    "plain:gfx950:256:16384:65536:2:0": [4, 512, 4096, 2, None],
    # 512 configurations, up to 8,192 charges, 713,262 lines, angle rule
    # (1,000 rows: 432 ms with 2,048-charge tiles, four workgroups per CU; 8
    # lanes per line within 1 %; 4,096-charge tiles, two workgroups per CU:
    # 452-467 ms; 8,192: 519 ms and more)
    "batch:gfx950:256:512:8192:1048576:2:0": [4, 256, 2048, 4, None],
}

TIE = 1.01  # timings within 1% of the best count as ties
UNROLLS = (8, 4, 2)  # second stage, in tie-break order


def _pick(report):
    """Fastest candidate of rows (S, block, tile, U, ms, ...); among ties, the
    earliest (fewest lanes per line; in the unroll stage, the larger U)."""
    best = min(r[4] for r in report)
    return next(r for r in report if r[4] <= best * TIE)


def _lookup(key: str):
    """Cached or shipped [S, block, tile, U, ms]; an entry written before the
    unroll was tuned ([S, block, tile, ms]) reads as U = 0 (library default)."""
    hit = _load().get(key) or SHIPPED.get(key)
    if hit and len(hit) == 4:
        hit = hit[:3] + [0] + hit[3:]
    return hit


def _launch(row):
    S, block, tile, U = row[:4]
    return api.make_launch(S, tile, block, U)


def _timed_row(S, block, tile, U, run):
    """(S, block, tile, U, ms), or ms = inf and the error for a shape the
    device cannot hold."""
    try:
        return (S, block, tile, U, run(api.make_launch(S, tile, block, U)))
    except Exception as e:
        return (S, block, tile, U, float("inf"), str(e)[:80])


def _tune(shapes, run):
    """Stage 1: every (S, block, tile) at the default unroll; stage 2: the
    fastest shape at each of UNROLLS.  Returns (best row, report)."""
    report = [_timed_row(S, block, tile, 0, run) for S, block, tile in shapes]
    first = _pick(report)
    if first[4] == float("inf"):
        return first, report
    second = [_timed_row(first[0], first[1], first[2], U, run) for U in UNROLLS]
    report += second
    ok = [r for r in second if r[4] < float("inf")]
    best = _pick(ok) if ok else first
    if best[4] * TIE > first[4]:
        best = first  # the default unroll stays unless an explicit one beats it by the margin
    return best, report


def _bucket(v: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, v)))))


def _round_up(a: int, b: int) -> int:
    return (a + b - 1) // b * b


def candidates(p: int, n: int, cus: int, batch: bool = False,
               stop: bool = False) -> List[Tuple[int, int, int]]:
    """(S, block, tile) shapes worth timing for p charges and n lines.

    Single configuration: S grows until the grid has 8 waves per SIMD.  Batch:
    S in {1, 2, 4, 8} always -- a smaller line block per work item evens out
    the persistent grid's finishing times even when waves are plentiful.
    With line ends (stop) and more than 16,384 charges the shape is that of
    the first phase of the lane-group trace, where a lane group takes a new
    line as soon as its line ends: S from 4 up to 32 (wide groups shorten the
    phase's last lines), blocks from one wave up, and smaller tiles as well."""
    out = []
    for S in (1, 2, 4, 8, 16, 32):
        if batch:
            if S > 8:
                continue
        elif stop:
            if S < 4:
                continue
        elif S > 1 and math.ceil(n * (S // 2) / 64) >= 8 * 4 * cus:
            continue  # already 8 waves per SIMD at S/2: splitting further only adds reductions
        if S > 1 and p / S < 32:
            continue
        for block in ((64, 128, 256, 512) if stop and not batch else (256, 512)):
            q = max(256, S * GROUP)
            need = _round_up(max(p, 1), q)
            tiles = set()
            if stop and not batch:
                tiles.update(t for t in (512, 1024, 2048, 4096) if t % q == 0 and t < need)
            for resident in ((1, 2, 4, 8) if batch else (1, 2, 4)):
                tmax = ((LDS_BYTES // resident - LDS_STATIC) // 16) // q * q
                if tmax < q:
                    continue
                if need <= tmax:
                    tiles.add(need)
                else:
                    k = math.ceil(p / tmax)
                    tiles.add(min(tmax, _round_up(math.ceil(p / k), q)))
            for t in sorted(tiles):
                out.append((S, block, t))
    return out


def _time(fn, reps: int = 3) -> float:
    import torch
    fn()  # warm-up (also JIT-free: kernels are precompiled)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune_plain(charges, seeds, rule: StepRule, rows: Optional[int] = None, use_cache: bool = True,
               max_candidates: Optional[int] = None):
    """Best launch for tracing ``seeds`` through ``charges`` with ``rule``.

    Candidates are timed on ``rows`` rows (48; 400 with line ends, so that the
    lines' different lengths show).  Returns (em_launch, report) where report
    lists (S, block, tile, U, ms); U = 0 is the library's default unroll."""
    import torch
    p, n = int(charges.shape[0]), int(seeds.shape[0])
    if not api.takes_launch_shape(p, rule):
        return api.make_launch(), [("register-resident line-ends trace: no launch shape",)]
    if rows is None:
        rows = 400 if rule.stop else 48
    key = f"plain:{_device_key()}:{_bucket(p)}:{_bucket(n)}:{rule.rule}:{int(rule.stop)}"
    if use_cache:
        hit = _lookup(key)
        if hit:
            return _launch(hit), [tuple(hit) + ("cached",)]
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    cands = candidates(p, n, cus, stop=rule.stop)
    if max_candidates:
        cands = cands[:max_candidates]
    buf = api.seed_rows(seeds, rows)
    best, report = _tune(cands, lambda launch: _time(
        lambda: api.trace_into(charges, buf, n, rule, launch)))
    _store(key, list(best[:5]))
    return _launch(best), report


def tune_batch(workloads, rule: StepRule, rows: int = 48, use_cache: bool = True):
    """Best launch for a batch plan over ``workloads`` (a list of gen.Workload):
    each candidate plan is timed on the same configurations with ``rows`` rows."""
    import torch
    ps = [w.p for w in workloads]
    ns = [w.n for w in workloads]
    key = (f"batch:{_device_key()}:{len(workloads)}:{_bucket(max(ps))}:{_bucket(sum(ns))}:"
           f"{rule.rule}:{int(rule.stop)}")
    if use_cache:
        hit = _lookup(key)
        if hit:
            return _launch(hit), [tuple(hit) + ("cached",)]
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    short = [type(w)(w.charges, w.seeds, rows) for w in workloads]
    layout, charges, lines = api.pack_batch(short)
    shapes = []
    for S, block, tile in candidates(max(ps), sum(ns), cus, batch=True):
        # persistent grid: several workgroups per CU keep every SIMD busy
        for t in sorted({tile, min(tile, _round_up(max(ps), max(256, S * GROUP)))}):
            if (S, block, t) not in shapes:
                shapes.append((S, block, t))

    def run(launch):
        plan = api.BatchPlan(layout, rule, launch, use_graph=False)
        try:
            return _time(lambda: plan.run(charges, lines))
        finally:
            plan.close()

    best, report = _tune(shapes, run)
    _store(key, list(best[:5]))
    return _launch(best), report
