"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs the reference's own CalcField_CPU<T> (ElectroMag/src/CPU_Implement.cpp,
compiled from /root/reference by ``make -C oracle ref`` into
oracle/_ref/libemref.so) on small seeded inputs and stores inputs and outputs
as compressed .npz files (no pickles).  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py

The fixtures are data: seeded charges and seeds (emtrace.gen), and the rows
the reference wrote.  The reference's process environment is the executable's
(-ffast-math links crtfastmath, which sets flush-to-zero); this script links
the reference without it, so denormals are kept -- irrelevant at these
magnitudes.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "electromag-with-cuda_amd"))

import oracle  # noqa: E402
from emtrace import gen  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# name: (p, n, half_side, len, resolution, use_curvature, dtype, seed)
CASES = {
    "ref_f32_plain_p64_n64_len200": (64, 64, 10.0, 200, 10.0, 0, np.float32, 0),
    "ref_f32_curv_p64_n64_len200": (64, 64, 10.0, 200, 10.0, 1, np.float32, 0),
    "ref_f64_plain_p64_n64_len200": (64, 64, 10.0, 200, 10.0, 0, np.float64, 0),
    "ref_f64_curv_p64_n64_len200": (64, 64, 10.0, 200, 10.0, 1, np.float64, 0),
    "ref_f32_curv_p1000_n128_len400": (1000, 128, 10.0, 400, 10.0, 1, np.float32, 0),
    "ref_f32_plain_p1000_n128_len400": (1000, 128, 10.0, 400, 10.0, 0, np.float32, 0),
    "ref_f32_curv_res1_p256_n64_len100": (256, 64, 5.0, 100, 1.0, 1, np.float32, 3),
    "ref_f32_curv_p1_n64_len50": (1, 64, 10.0, 50, 10.0, 1, np.float32, 7),
    "ref_f32_plain_p1_n64_len50": (1, 64, 10.0, 50, 10.0, 0, np.float32, 7),
    "ref_f32_curv_p64_n64_len2": (64, 64, 10.0, 2, 10.0, 1, np.float32, 5),
}

# argument errors the reference reports (CPU_Implement.cpp:43-52, 213-230)
ERROR_CASES = {
    # name: (n, len, resolution, use_curvature, p)
    "n0": (0, 10, 10.0, 1, 4),
    "res0_curv": (64, 10, 0.0, 1, 4),
    "res0_plain": (64, 10, 0.0, 0, 4),
    "len1_curv": (64, 1, 10.0, 1, 4),
    "len1_plain": (64, 1, 10.0, 0, 4),
    "n_not_64_curv": (32, 10, 10.0, 1, 4),
    "n_not_64_plain": (32, 10, 10.0, 0, 4),
}


def main() -> None:
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libemref.so missing: run `make -C oracle ref` first")
    for name, (p, n, half, length, res, uc, dt, seed) in CASES.items():
        w = gen.plain_config(p, n, half, length, seed)
        rc, rows, _, _ = oracle.ref_calc_lines(w.charges, w.seeds, length, res, uc, dtype=dt)
        assert rc == 0, (name, rc)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), charges=w.charges.astype(dt),
                            seeds=w.seeds.astype(dt), rows=rows, resolution=np.array(res),
                            use_curvature=np.array(uc), rc=np.array(rc))
        print(f"{name}: rows {rows.shape} {rows.dtype}, nan lines "
              f"{int(np.isnan(rows).any(axis=(0, 1)).sum())}")
    codes = {}
    for name, (n, length, res, uc, p) in ERROR_CASES.items():
        ch = np.zeros((p, 4), np.float32)
        seeds = np.zeros((n, 3), np.float32)
        rc, _, _, _ = oracle.ref_calc_lines(ch, seeds, length, res, uc) if n else \
            oracle.ref_calc_lines(ch, np.zeros((1, 3), np.float32)[:0], length, res, uc)
        codes[name] = rc
    np.savez_compressed(os.path.join(OUT, "ref_error_codes.npz"),
                        names=np.array(list(codes.keys())),
                        codes=np.array(list(codes.values()), dtype=np.int32),
                        args=np.array([ERROR_CASES[k][:4] for k in codes], dtype=np.float64))
    print("error codes:", codes)
    legacy()


# The reference application's own problems (ElectroMag.cpp:43-58 presets,
# Electromag utils.h:29-111 initialisation), traced by the reference's
# CalcField_CPU<float> with the application's arguments (resolution 1,
# useCurvature true; ElectroMag.cpp:296, 305-306).  Lines are independent, so
# a subset of a preset's seeds traces to the same rows as the full run.
LEGACY_GRIDS = [(16, 16, 1), (5, 3, 4), (7, 1, 3), (128, 128, 1)]
LEGACY_TRACES = {
    # name: (nx, ny, nz, p, len, line stride, row stride)
    "legacy_bogo_ref": (16, 16, 1, 50, 500, 4, 2),
    "legacy_default_ref": (128, 128, 1, 1024, 2500, 256, 10),
}


def legacy() -> None:
    out = {"charges_p1024": oracle.ref_init_charges(1024)}
    for (w, h, d) in LEGACY_GRIDS:
        out[f"grid_{w}x{h}x{d}"] = oracle.ref_init_grid(w, h, d)
    np.savez_compressed(os.path.join(OUT, "legacy_init.npz"), **out)
    print("legacy_init:", {k: v.shape for k, v in out.items()})
    for name, (nx, ny, nz, p, length, ls, rs) in LEGACY_TRACES.items():
        charges = oracle.ref_init_charges(p)
        seeds = oracle.ref_init_grid(nx, ny, nz)[::ls]
        assert seeds.shape[0] % 64 == 0  # the SSE path's line multiple
        rc, rows, _, _ = oracle.ref_calc_lines(charges, seeds, length, 1.0, 1)
        assert rc == 0, (name, rc)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), lines=np.arange(0, nx * ny * nz, ls),
                            row_stride=np.array(rs), rows=rows[:, ::rs, :],
                            preset=np.array([nx, ny, nz, p, length]))
        print(f"{name}: lines {seeds.shape[0]}, rows {rows[:, ::rs, :].shape}")


if __name__ == "__main__":
    main()
