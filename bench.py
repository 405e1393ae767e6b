#!/usr/bin/env python3
"""Benchmark of the gfx950 field-line tracer.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A "step" is one full trace of the headline workload: 16,384 charges in
[-50,50]^3, 65,536 lines, 2,500 rows, curvature-adaptive (angle) step in
[0.005, 0.5], every line advanced all 2,499 steps (charge-point interactions
= p * n * (len-1) = 2.683e12 per step).  Inputs are resident in HBM before the
timed region; the kernel time comes from HIP events on the stream the kernel
runs on.  Prints one JSON line.

Roofline: 18 credited FLOP per interaction (15 for the Coulomb term + 3 to
accumulate, the reference's accounting: include/Electrostatics.h:62,
ElectroMag/src/CPU_Implement.cpp:32) against MI355X's 157.3 TFLOP/s FP32 peak.
The CPU baseline times the reference's own CalcField_CPU<float> (compiled from
source into oracle/_ref), or the C restatement in oracle/ where that build is
absent, on a bounded slice of the same workload.

Extra fields (skipped with --no-extras): the same workload with line ends at
sinks (useful interactions only), lines_8k in full on the GPU and through the
reference's CPU path, the batched workload (512 configurations,
one hipGraph-replayed persistent launch), the differentiation workload
(forward + adjoint time over forward time), the inverse fit, the FP64 kernel,
the reference application's default problem and the particle step.
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
    ap.add_argument("--workload", default="lines_64k", choices=["lines_64k", "lines_8k"])
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--unroll", type=int, default=0, help="charge records per pipelined group (2, 4, 8)")
    ap.add_argument("--no-tune", action="store_true", help="use the library's heuristic shape")
    ap.add_argument("--cpu-lines", type=int, default=1024,
                    help="lines of the workload timed on the CPU (multiple of 64)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="all",
                    help="comma-separated extra workloads to run (default: all): lines_64k_with_line_ends, "
                         "lines_8k, batch_512, adjoint_2k, fit_64, fit_well_posed, fp64_kernel, "
                         "legacy_default_preset, particle_step")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run a 1 GiB device copy (calibrates the FETCH/WRITE_SIZE counters)")
    return ap.parse_args()


def timed(fn, steps, warmup):
    """Run fn warmup times, then time `steps` runs: (wall s, [kernel s each])."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [a.elapsed_time(b) * 1e-3 for a, b in evs]


def host_cpu():
    """The host CPU as the baseline runs on it: model name with its x86 family,
    model and stepping numbers (a virtual machine may report only a family
    name), the SIMD extensions the reference's SSE path could use, CPUs this
    process may run on (affinity), CPUs in the machine, and the OpenMP thread
    count a parallel region gets (what the reference's OpenMP loop uses)."""
    import ctypes
    model, ident, flags = "unknown", {}, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model == "unknown":
                    model = val
                elif key in ("vendor_id", "cpu family", "model", "stepping") and key not in ident:
                    ident[key] = val
                elif key == "flags" and not flags:
                    flags = set(val.split())
                elif not line.strip() and model != "unknown":
                    break  # the first processor's block is enough
    except OSError:
        pass
    if ident:
        model += " (" + ", ".join(f"{k} {v}" for k, v in ident.items()) + ")"
    try:
        omp = int(ctypes.CDLL("libgomp.so.1").omp_get_max_threads())
    except OSError:
        omp = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    affinity = len(os.sched_getaffinity(0))
    simd = [f for f in ("sse2", "avx2", "avx512f") if f in flags]
    info = {"cpu_model": model, "cpu_simd": simd, "omp_threads": omp, "affinity_cpus": affinity,
            "machine_cpus": os.cpu_count()}
    if omp < affinity:
        info["threads_note"] = (f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')} caps the "
                                f"OpenMP team below the {affinity} CPUs this process may use")
    return info


def cpu_baseline(w, lines):
    """Time the reference (or the restatement) on `lines` lines of the workload."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import oracle  # the checker / CPU baseline only
    from emtrace import rules
    host = host_cpu()
    threads = host["omp_threads"]
    lines = max(64, lines // 64 * 64)
    seeds = w.seeds[np.linspace(0, w.n - 1, lines).astype(np.int64)]
    t0 = time.perf_counter()
    if oracle.ref_lib() is not None:
        rc, _, _, _ = oracle.ref_calc_lines(w.charges, seeds, w.len, 10.0, 1)
        kind = "reference"
        path = "reference CalcField_CPU<float>, useCurvature=true (SSE + OpenMP)"
    else:
        rc, _ = oracle.calc_lines(w.charges, seeds, w.len, rules.ref(10.0))
        kind = "port"
        path = "oracle/restate.c (scalar C + OpenMP)"
    dt = time.perf_counter() - t0
    if rc:
        return None
    inter = w.p * lines * (w.len - 1)
    return {"value": inter / dt, "unit": "interactions/s", "cores": threads, "kind": kind,
            "sample": f"{path}: {w.p} charges x {lines} of {w.n} lines x {w.len} rows "
                      f"(every line all {w.len - 1} steps, step 0.1), {dt:.2f} s wall; cost is "
                      f"linear in the line count", "seconds": dt, **host}


def copy_only():
    import torch
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import emtrace
    from emtrace import autotune, gen, rules

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    torch.cuda.init()
    emtrace._lib.check(emtrace.lib().em_init(0), "em_init")
    if args.copy_only:
        return copy_only()

    rule = rules.angle(stop=False)
    w = gen.lines_64k() if args.workload == "lines_64k" else gen.lines_8k()
    charges = torch.from_numpy(w.charges).cuda()
    seeds = torch.from_numpy(w.seeds).cuda()
    if args.lanes or args.tile or args.block or args.unroll:
        launch, tuned = emtrace.make_launch(args.lanes, args.tile, args.block, args.unroll), "given"
    elif args.no_tune:
        launch, tuned = None, "heuristic"
    else:
        launch, report = autotune.tune_plain(charges, seeds, rule)
        tuned = "autotuned (cache or shipped table)" if report[0][-1] == "cached" else "autotuned"
    rows = emtrace.seed_rows(seeds, w.len)
    steps_buf = torch.empty(w.n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    wall, kern = timed(lambda: emtrace.trace_into(charges, rows, w.n, rule, launch, steps=steps_buf),
                       args.steps, args.warmup)
    k_avg = sum(kern) / len(kern)
    inter = float(w.interactions)
    value = inter * args.steps / wall
    achieved = inter * FLOP_PER_INTERACTION / k_avg / 1e12
    finite = bool(torch.isfinite(rows).all().item())
    all_steps = bool((steps_buf == w.len - 1).all().item())
    shape = ((launch.lanes_per_line, launch.tile, launch.block, launch.unroll)
             if launch is not None else None)

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
        "config": {"workload": f"{args.workload}: plain method, {w.p} charges in "
                               f"[-{50 if args.workload == 'lines_64k' else 10}, ...]^3, {w.n} lines, "
                               f"{w.len} rows, curvature-adaptive (angle) step in [0.005, 0.5], "
                               f"every line all {w.len - 1} steps",
                   "charges": w.p, "lines": w.n, "rows": w.len, "step_rule": "angle",
                   "launch": {"lanes_per_line": shape[0], "tile": shape[1], "block": shape[2],
                              "unroll": shape[3], "source": tuned} if shape else {"source": tuned},
                   "parallelism": "single GPU"},
        "line_steps_per_s": w.n * (w.len - 1) * args.steps / wall,
        "frac_fp32_peak": value * FLOP_PER_INTERACTION / (PEAK_FP32_TFLOPS * 1e12),
        "finite_output": finite,
        "all_lines_full_length": all_steps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "peak_source": "MI355X FP32 vector peak (157.3 TFLOP/s, the same figure as dense "
                                    "FP32 MFMA); the kernel is bound by FP32 VALU issue, neither by "
                                    "HBM (1.97 GB per launch) nor by MFMA (K = 3 is no contraction)",
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel_ms": k_avg * 1e3, "flop_per_launch": inter * FLOP_PER_INTERACTION,
                     "algorithmic_bytes_per_launch": 12.0 * w.n * (w.len - 1)},
    }
    hw_flop = profiled_value("trace_lines64k_valu_counters.md", "fp32_flop_per_launch",
                             args.workload, shape)
    if hw_flop is not None:
        out["roofline"]["rocprof_fp32"] = {
            "flop_per_launch": hw_flop, "achieved": hw_flop / k_avg / 1e12,
            "frac": hw_flop / k_avg / 1e12 / PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "source": "profiles/trace_lines64k_valu_counters.md (rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 "
                      "x 64 per dispatch) over this run's kernel time"}
    traffic = profiled_traffic(args.workload, shape)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_source"] = "profiles/trace_lines64k_hbm_traffic.md (rocprofv3 " \
            "FETCH_SIZE x2.00 + WRITE_SIZE x1.00, scales from a 1 GiB copy)"
    if not args.no_extras:
        only = None if args.extras == "all" else set(args.extras.split(","))
        out["extras"] = extras(w, charges, seeds, value, launch, only)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(w, args.cpu_lines)
    print(json.dumps(out), flush=True)


def profiled_value(name, key, workload, shape):
    """A per-launch figure from a committed counter profile, when this run used
    the launch shape that was profiled (lanes 4, tile 4096, block 512, unroll 2)."""
    path = os.path.join(ROOT, "profiles", name)
    if workload != "lines_64k" or shape != (4, 4096, 512, 2) or not os.path.exists(path):
        return None
    for line in open(path):
        if f"{key}:" in line:
            return float(line.split(f"{key}:")[1].split("*")[0].strip(" *\n"))
    return None


def profiled_traffic(workload, shape):
    """HBM bytes per launch (calibrated FETCH_SIZE + WRITE_SIZE)."""
    return profiled_value("trace_lines64k_hbm_traffic.md", "traffic_bytes_per_launch", workload,
                          shape)


def extras(w, charges, seeds, plain_value, launch, only=None):
    """The extra workloads (all, or those named in ``only``); returns their records."""
    import torch
    import emtrace
    from emtrace import autotune, gen, rules
    res = {}

    def on(name):
        return only is None or name in only

    # 1. same workload, lines end where they reverse (useful interactions only)
    if on("lines_64k_with_line_ends"):
        rule_s = rules.angle(stop=True)
        rows = emtrace.seed_rows(seeds, w.len)
        st = torch.empty(w.n, dtype=torch.int32, device="cuda")
        wall, kern = timed(lambda: emtrace.trace_into(charges, rows, w.n, rule_s, steps=st), 3, 1)
        taken = float(st.double().sum().item())
        best = min(kern)
        res["lines_64k_with_line_ends"] = {
            "useful_interactions_per_s": w.p * taken / best, "kernel_ms": best * 1e3,
            "kernel_ms_each": [k * 1e3 for k in kern],
            "mean_steps": taken / w.n,
            "lines_over_192_steps": int((st > 192).sum().item()),
            "note": "lines end at sinks; the register-resident trace: one 1,024-lane workgroup "
                    "per CU holds all 16,384 charge records in registers and traces up to 16 "
                    "lines at once in two pipelined groups; lines come from a counter until it "
                    "runs out, then every workgroup's lines are settled for up to 128 steps, "
                    "then the lines still moving are dealt out by work left and traced to the "
                    "end, while that last pass's idle waves write the rows after the end of "
                    "every line that ended before it. Best of 3 calls (kernel sequence "
                    "event-timed)"}
        del rows
    # 1b. lines_8k (1,000 charges in [-10,10]^3, 8,192 lines, 2,500 rows) in full:
    #     the GPU kernel (ref rule, resolution 10: the reference's live default
    #     step; and the angle rule, every line all steps) and the reference's
    #     own CalcField_CPU<float> with useCurvature on all 8,192 lines
    if on("lines_8k"):
        w0 = gen.lines_8k()
        c0, s0 = torch.from_numpy(w0.charges).cuda(), torch.from_numpy(w0.seeds).cuda()
        r0 = emtrace.seed_rows(s0, w0.len)
        rec = {"charges": w0.p, "lines": w0.n, "rows": w0.len, "interactions": float(w0.interactions)}
        for name, rule0 in (("ref", rules.ref(10.0)), ("angle", rules.angle(stop=False))):
            _, k0 = timed(lambda: emtrace.trace_into(c0, r0, w0.n, rule0), 3, 1)
            rec[f"gpu_{name}_kernel_ms"] = min(k0) * 1e3
            rec[f"gpu_{name}_interactions_per_s"] = float(w0.interactions) / min(k0)
        del r0
        sys.path.insert(0, ROOT)
        import oracle  # the CPU baseline only
        if oracle.ref_lib() is not None:
            t0 = time.perf_counter()
            rc, _, t_in, _ = oracle.ref_calc_lines(w0.charges, w0.seeds, w0.len, 10.0, 1)
            dt = time.perf_counter() - t0
            if rc == 0:
                rec["cpu_reference_s"] = dt
                rec["cpu_reference_interactions_per_s"] = float(w0.interactions) / dt
                rec["gpu_ref_speedup"] = dt / (rec["gpu_ref_kernel_ms"] * 1e-3)
                rec["cpu_note"] = ("reference CalcField_CPU<float>, useCurvature=true (SSE + "
                                   "OpenMP), all 8,192 lines, wall time of the call")
                rec.update({k: v for k, v in host_cpu().items()})
        res["lines_8k"] = rec
    # 2. batched workload (512 configurations, 1,000 rows), one graph-replayed launch
    if on("batch_512"):
        ws = gen.batch_512()
        rule_b = rules.angle(stop=False)
        blaunch, _ = autotune.tune_batch(ws, rule_b)
        layout, bch, blines = emtrace.pack_batch(ws)
        plan = emtrace.BatchPlan(layout, rule_b, blaunch, use_graph=True)
        wall, kern = timed(lambda: plan.run(bch, blines), 2, 1)
        binter = float(layout.interactions())
        bval = binter * 2 / wall
        info = plan.info()
        res["batch_512"] = {
            "interactions": binter, "interactions_per_s": bval, "kernel_ms": min(kern) * 1e3,
            "ratio_to_plain": bval / plain_value, "configs": len(ws), "items": info["items"],
            "grid": info["grid"], "launch": {k: info[k] for k in ("lanes_per_line", "tile", "block", "unroll")}}
        plan.close()
        del blines, bch
    # 3. differentiation workload: forward vs forward + adjoint
    if on("adjoint_2k"):
        w3 = gen.adjoint_2k()
        tq = gen.perturbed(w3.charges, 0.05)
        c3 = torch.from_numpy(w3.charges).cuda()
        s3 = torch.from_numpy(w3.seeds).cuda()
        rule_d = rules.angle()
        tgt = emtrace.trace(torch.from_numpy(tq).cuda(), s3, w3.len, rule_d)
        every = 50
        rows3 = emtrace.seed_rows(s3, w3.len)
        ck = torch.empty(emtrace.ckpt_size(w3.n, w3.len, every), dtype=torch.float32, device="cuda")

        def fwd():
            emtrace.trace_into(c3, rows3, w3.n, rule_d, None, ck, every)

        def fwd_adj():
            fwd()
            g = (2.0 / (w3.n * w3.len)) * (rows3 - tgt)
            emtrace.adjoint(c3, rows3, ck, every, g, rule_d)

        _, kf = timed(fwd, 5, 2)
        _, kfa = timed(fwd_adj, 5, 2)
        res["adjoint_2k"] = {"forward_ms": min(kf) * 1e3, "forward_adjoint_ms": min(kfa) * 1e3,
                                  "ratio": min(kfa) / min(kf), "ckpt_every": every}
        del rows3, ck, tgt
    # 4. inverse fit: 64 configurations, 800 rows, 200 Adam iterations, each one
    #    captured batched forward + adjoint + update
    if on("fit_64"):
        from emtrace.fit import InverseFit
        ws4 = gen.fit_64()
        truth = [gen.perturbed(w4.charges, 0.1, seed=i) for i, w4 in enumerate(ws4)]
        fit = InverseFit(ws4, truth, rules.angle(), ckpt_every=50)
        fit.plan.run(fit.theta, fit.lines)
        loss0 = fit.loss()
        fit.step()  # eager warm-up iteration + graph capture
        torch.cuda.synchronize()
        iters = 199
        t0 = time.perf_counter()
        for _ in range(iters):
            fit.step()
        torch.cuda.synchronize()
        per_iter = (time.perf_counter() - t0) / iters
        fit.plan.run(fit.theta, fit.lines)
        res["fit_64"] = {
            "configs": len(ws4), "charges": int(fit.layout.p.sum()), "lines": int(fit.layout.n.sum()),
            "iterations": iters + 1, "ms_per_iteration": per_iter * 1e3, "loss_start": loss0,
            "loss_end": fit.loss(), "recovery": fit.recovery(),
            "note": "iteration = batched forward with checkpoints + dL/dx + batched adjoint + Adam, "
                    "replayed from one captured graph"}
        del fit
    # 4b. a well-posed inverse problem through the same machinery: 64
    #     configurations of 2 charges, 256 lines of 60 rows, positions perturbed
    #     by N(0, 0.02^2), plain rule, 200 iterations with a cosine-decayed rate
    if on("fit_well_posed"):
        import math
        from emtrace.fit import InverseFit
        wsw = gen.batched_configs(64, (2, 2), (256, 256), 60, 5)
        truthw = [gen.perturbed(w4.charges, 0.02, seed=100 + i) for i, w4 in enumerate(wsw)]
        fitw = InverseFit(wsw, truthw, rules.plain(10.0), ckpt_every=20, lr_pos=4e-3, lr_q=2e-3)
        fitw.plan.run(fitw.theta, fitw.lines)
        lw0 = fitw.loss()
        t0 = time.perf_counter()
        for it in range(200):
            fitw.set_lr_scale(0.5 * (1.0 + math.cos(math.pi * it / 200)))
            fitw.step()
        torch.cuda.synchronize()
        per_w = (time.perf_counter() - t0) / 200
        fitw.plan.run(fitw.theta, fitw.lines)
        res["fit_well_posed"] = {
            "configs": len(wsw), "charges_per_config": 2, "lines_per_config": 256, "rows": 60,
            "sigma": 0.02, "rule": "plain", "iterations": 200, "ms_per_iteration": per_w * 1e3,
            "loss_start": lw0, "loss_end": fitw.loss(), "recovery": fitw.recovery(),
            "note": "same forward + adjoint + Adam graph as fit_64; learning rate 4e-3 (positions) / "
                    "2e-3 (magnitudes), cosine-decayed; magnitudes compared modulo each "
                    "configuration's overall scale (q_rms_scaled), which field lines cannot see"}
        del fitw
    # 5. FP64 kernel (IEEE binary64, reference op order): lines_64k charges and
    #    lines, 250 rows, plain step
    if on("fp64_kernel"):
        rows_n = 250
        q64, s64 = charges.double(), seeds.double()
        rule_p = rules.plain(10.0)
        r64 = torch.empty((3, rows_n, w.n), dtype=torch.float64, device="cuda")
        r64[:, 0, :] = s64.t()

        def fp64():
            emtrace._lib.call("em_trace_dev_f64", q64.data_ptr(), w.p, r64[0].data_ptr(),
                              r64[1].data_ptr(), r64[2].data_ptr(), w.n, rows_n, w.n,
                              emtrace.api.ctypes.byref(emtrace.make_step(rule_p)), None,
                              emtrace.api._stream_handle())

        _, k64 = timed(fp64, 2, 1)
        i64 = float(w.p) * w.n * (rows_n - 1)
        res["fp64_kernel"] = {"interactions_per_s": i64 / min(k64), "kernel_ms": min(k64) * 1e3,
                              "rows": rows_n, "step_rule": "plain",
                              "frac_fp64_peak": i64 / min(k64) * FLOP_PER_INTERACTION / 78.6e12,
                              "note": "one lane per line, charges summed in index order; a correctly "
                                      "rounded sqrt and division per interaction (plain rule)"}
        del r64, q64, s64
    # 6. the reference application's default problem (ElectroMag.cpp:43: 1,024
    #    charges, 128 x 128 grid seeds, 2,500 rows, curvature step, resolution 1)
    if on("legacy_default_preset"):
        wl = emtrace.legacy.preset("default")
        cl, sl = torch.from_numpy(wl.charges).cuda(), torch.from_numpy(wl.seeds).cuda()
        rl = emtrace.seed_rows(sl, wl.len)
        _, kl = timed(lambda: emtrace.trace_into(cl, rl, wl.n, rules.ref(1.0)), 3, 1)
        # the reference application's own call (ElectroMag.cpp:305): host arrays
        # freshly allocated (pages not yet touched), through the drop-in entry
        import numpy as np
        calls = []
        for _ in range(2):
            hx, hy, hz = (np.empty(wl.len * wl.n, np.float32) for _ in range(3))
            hx[:wl.n], hy[:wl.n], hz[:wl.n] = wl.seeds[:, 0], wl.seeds[:, 1], wl.seeds[:, 2]
            t0 = time.perf_counter()
            rc, perf = emtrace.calc_lines(wl.charges, hx, hy, hz, wl.n, 1.0, True)
            calls.append((time.perf_counter() - t0, perf))
            if rc != 0:
                raise SystemExit(f"drop-in call failed: {rc}")
            del hx, hy, hz
        wall_d, perf_d = min(calls, key=lambda c: c[0])
        # the same call on doubles (CalcField_GPU<double> -> em_calc_lines_f64)
        calls64 = []
        for _ in range(2):
            hx, hy, hz = (np.empty(wl.len * wl.n, np.float64) for _ in range(3))
            hx[:wl.n], hy[:wl.n], hz[:wl.n] = wl.seeds[:, 0], wl.seeds[:, 1], wl.seeds[:, 2]
            t0 = time.perf_counter()
            rc, perf = emtrace.calc_lines64(wl.charges.astype(np.float64), hx, hy, hz, wl.n, 1.0, True)
            calls64.append((time.perf_counter() - t0, perf))
            if rc != 0:
                raise SystemExit(f"FP64 drop-in call failed: {rc}")
            del hx, hy, hz
        wall_64, perf_64 = min(calls64, key=lambda c: c[0])
        res["legacy_default_preset"] = {
            "interactions_per_s": float(wl.interactions) / min(kl), "kernel_ms": min(kl) * 1e3,
            "charges": wl.p, "lines": wl.n, "rows": wl.len,
            "dropin_host_arrays": {
                "call_s": wall_d, "h2d_s": perf_d["h2d_s"], "kernel_s": perf_d["kernel_s"],
                "d2h_s": perf_d["d2h_s"], "d2h_GB_per_s": perf_d["d2h_bytes"] / perf_d["d2h_s"] / 1e9,
                "reference_gflops_formula": perf_d["gflops"],
                "note": "em_calc_lines_f32 on fresh numpy arrays (best of 2 calls); the reference "
                        "app's CalcField call. Rows are traced in row chunks; each chunk's copy "
                        "back into the pageable arrays (d2h_s sums the transfers) overlaps the "
                        "next chunk's kernel"},
            "dropin_host_arrays_f64": {
                "call_s": wall_64, "h2d_s": perf_64["h2d_s"], "kernel_s": perf_64["kernel_s"],
                "d2h_s": perf_64["d2h_s"],
                "d2h_GB_per_s": perf_64["d2h_bytes"] / perf_64["d2h_s"] / 1e9,
                "interactions_per_s": perf_64["interactions_per_s"],
                "note": "em_calc_lines_f64 (the IEEE binary64 kernel) on fresh float64 arrays, best "
                        "of 2 calls; cached device buffers, row chunks as the FP32 call"}}
        del rl, cl, sl
    # 7. particle step (Particle_System.cpp:108-190): 65,536 particles among the
    #    16,384 lines_64k charges, one step with the written-back state
    if on("particle_step"):
        import numpy as np
        rng = np.random.default_rng(0)
        npart = 65536
        P = np.zeros((npart, 8), np.float32)
        P[:, :3] = rng.uniform(-50, 50, (npart, 3))
        P[:, 3] = rng.uniform(-1e-3, 1e-3, npart)
        P[:, 4:7] = rng.normal(0, 1, (npart, 3))
        P[:, 7] = 1.0
        dP = torch.from_numpy(P).cuda()
        dA = torch.zeros((npart, 3), dtype=torch.float32, device="cuda")
        dO = torch.empty_like(dP)
        _, kv = timed(lambda: emtrace.verlet_step_dev(dP, charges, dA, [-50] * 3, [50] * 3, 1e-3, out=dO),
                      3, 1)
        pairs = float(npart) * (w.p + npart - 1)
        res["particle_step"] = {"particles": npart, "static_charges": w.p, "interactions": pairs,
                                "interactions_per_s": pairs / min(kv), "kernel_ms": min(kv) * 1e3,
                                "note": "static terms 12 VALU + rsq, particle-particle terms (Coulomb + "
                                        "Biot-Savart, self masked) 22 VALU + rsq"}
    return res

if __name__ == "__main__":
    main()
