#!/usr/bin/env python3
"""Benchmark of the gfx950 field-line tracer.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config1|config0|config2]

A "step" is one full trace of the workload: every line advanced len-1 steps
through the O(p) Coulomb sum (charge-point interactions = p * n * (len-1)).
Default workload: 16,384 charges in [-50,50]^3, 65,536 lines, 2,500 rows,
curvature-adaptive (angle) step in [0.005, 0.5] -- inputs resident in HBM
before the timed region.  Prints one JSON line.

The roofline figure uses 18 credited FLOP per interaction (15 for the Coulomb
term + 3 to accumulate, the accounting of the reference's
include/Electrostatics.h:62 and ElectroMag/src/CPU_Implement.cpp:32) against
MI355X's 157.3 TFLOP/s FP32 peak; the kernel time comes from HIP events
recorded on the stream the kernel runs on.

The CPU baseline times the reference's own CalcField_CPU<float> (compiled
from source into oracle/_ref) -- or, where that build is absent, the C
restatement in oracle/ -- on a bounded slice of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "electromag-with-cuda_amd"))

METRIC = "charge-point interactions/s (% FP32 peak) @16k q×64k lines; batched & adj/fwd"
PEAK_FP32_TFLOPS = 157.3
FLOP_PER_INTERACTION = 18


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="config1", choices=["config1", "config0", "config2"])
    ap.add_argument("--rule", default="angle", choices=["angle", "ref", "plain"])
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-lines", type=int, default=1024,
                    help="lines of the workload timed on the CPU (multiple of 64)")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def cpu_baseline(w, rule, lines):
    """Time the reference (or the restatement) on `lines` lines of the workload."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import oracle  # the checker / CPU baseline only
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    lines = max(64, lines // 64 * 64)
    seeds = w.seeds[np.linspace(0, w.n - 1, lines).astype(np.int64)]
    res = 1.0 / rule.h0 if rule.rule == 2 else rule.resolution
    t0 = time.perf_counter()
    if oracle.ref_lib() is not None:
        rc, _, t_ref, _ = oracle.ref_calc_lines(w.charges, seeds, w.len, res, 1)
        kind = "reference"
        path = "CalcField_CPU<float>, useCurvature=true (SSE + OpenMP)"
    else:
        rc, _ = oracle.calc_lines(w.charges, seeds, w.len, oracle_rule(rule))
        kind = "port"
        path = "oracle/restate.c (scalar C + OpenMP)"
    dt = time.perf_counter() - t0
    if rc:
        return None
    inter = w.p * lines * (w.len - 1)
    return {"value": inter / dt, "unit": "interactions/s", "cores": threads, "kind": kind,
            "sample": f"{path}: {w.p} charges x {lines} of {w.n} lines x {w.len} rows, "
                      f"step 1/{res:g}, {dt:.2f} s wall; cost is linear in the line count",
            "seconds": dt}


def oracle_rule(rule):
    return rule


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import emtrace
    from emtrace import gen, rules

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    torch.cuda.init()
    L = emtrace.lib()
    emtrace._lib.check(L.em_init(0), "em_init")
    rule = {"angle": rules.angle(), "ref": rules.ref(10.0), "plain": rules.plain(10.0)}[args.rule]
    launch = emtrace.make_launch(args.lanes, args.tile, args.block)

    if args.workload == "config2":
        return bench_batched(args, rule, launch)

    w = gen.config1() if args.workload == "config1" else gen.config0()
    charges = torch.from_numpy(w.charges).cuda()
    seeds = torch.from_numpy(w.seeds).cuda()
    rows = emtrace.seed_rows(seeds, w.len)
    torch.cuda.synchronize()

    def step():
        emtrace.trace_into(charges, rows, w.n, rule, launch)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = [a.elapsed_time(b) * 1e-3 for a, b in evs]
    k_avg = sum(kern) / len(kern)
    inter = float(w.interactions)
    value = inter * args.steps / wall
    achieved_tflops = inter * FLOP_PER_INTERACTION / k_avg / 1e12
    ok = bool(torch.isfinite(rows).all().item())

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, rule, args.cpu_lines)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: plain method, {w.p} charges, {w.n} lines, "
                               f"{w.len} rows, {rules.NAMES[rule.rule]} step rule",
                   "charges": w.p, "lines": w.n, "rows": w.len, "step_rule": rules.NAMES[rule.rule],
                   "parallelism": "single GPU"},
        "line_steps_per_s": w.n * (w.len - 1) * args.steps / wall,
        "frac_fp32_peak": value * FLOP_PER_INTERACTION / (PEAK_FP32_TFLOPS * 1e12),
        "finite_output": ok,
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP32_TFLOPS,
                     "traffic": None, "kernel_ms": k_avg * 1e3,
                     "flop_per_launch": inter * FLOP_PER_INTERACTION},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_batched(args, rule, launch):
    import torch
    import emtrace
    from emtrace import gen
    ws = gen.config2()
    layout, charges, lines = emtrace.pack_batch(ws)
    plan = emtrace.BatchPlan(layout, rule, launch, use_graph=True)
    info = plan.info()
    for _ in range(args.warmup):
        plan.run(charges, lines)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        plan.run(charges, lines)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k_avg = sum(a.elapsed_time(b) for a, b in evs) / len(evs) * 1e-3
    inter = float(layout.interactions())
    value = inter * args.steps / wall
    out = {"metric": "batched charge-point interactions/s (config 2)", "value": value,
           "unit": "interactions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"config2: {len(ws)} configurations, {int(layout.p.sum())} charges, "
                                  f"{int(layout.n.sum())} lines, 1000 rows, hipGraph persistent launch",
                      "plan": info},
           "frac_fp32_peak": value * FLOP_PER_INTERACTION / (PEAK_FP32_TFLOPS * 1e12),
           "kernel_ms": k_avg * 1e3}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
