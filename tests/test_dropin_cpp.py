"""The C++ drop-in (include/emtrace/calcfield.hpp) at the reference's call site.

CPU: the header compiles against the reference's own containers
(Vector3<Array<float>>, Array<pointCharge<float>>, perfPacket from
include/SOA_utils.hpp, include/Data Structures.h, include/Electrostatics.h)
in the call shape of ElectroMag.cpp:305-306, and against its standalone
compat types with -std=c++11 (the reference builds with -std=c++0x).
GPU: the compiled drop-in program writes the same rows as the device path.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "electromag-with-cuda_amd", "lib")
MAIN = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_main")
REFERENCE = "/root/reference"

CALL_SITE = r'''
#include "CPU Implement.h"
#include "emtrace/calcfield.hpp"
int run(size_t n, size_t len, size_t p) {
    Vector3<Array<float> > CPUlines;
    Array<electro::pointCharge<float> > charges;
    CPUlines.AlignAlloc(n * len);
    charges.AlignAlloc(p);
    perfPacket CPUperf;
    return emtrace::CalcField_GPU(CPUlines, charges, n, 1.0f, CPUperf, true);
}
int run_aos(size_t n, size_t len, size_t p) {  // the GPU stub's shape (Electrostatics.cpp:28-42)
    Array<Vector3<float> > lines;
    Array<electro::pointCharge<float> > charges;
    lines.AlignAlloc(n * len);
    charges.AlignAlloc(p);
    perfPacket perf;
    return emtrace::CalcField(lines, charges, n, 1.0f, perf, true);
}
int run_f64(size_t n, size_t len, size_t p) {  // CalcField_CPU<double>'s containers
    Vector3<Array<double> > lines;
    Array<electro::pointCharge<double> > charges;
    lines.AlignAlloc(n * len);
    charges.AlignAlloc(p);
    perfPacket perf;
    return emtrace::CalcField_GPU(lines, charges, n, 1.0, perf, true);
}
'''


def build_dropin() -> str:
    if not os.path.exists(BIN):
        os.makedirs(os.path.dirname(BIN), exist_ok=True)
        subprocess.run(["g++", "-std=c++11", "-O2", "-I", INC, MAIN, "-L", LIBDIR, "-lemtrace",
                        f"-Wl,-rpath,{LIBDIR}", "-o", BIN], check=True)
    return BIN


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree absent")
def test_header_accepts_reference_types():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "site.cpp")
        open(src, "w").write(CALL_SITE)
        subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-w", "-I", INC,
                        "-I", os.path.join(REFERENCE, "include"),
                        "-I", os.path.join(REFERENCE, "ElectroMag", "src"), src], check=True)


def test_header_and_program_compile_standalone():
    with tempfile.TemporaryDirectory() as d:
        subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I", INC, MAIN],
                       check=True)


def _run_dropin(charges, seeds, length, resolution, curvature, aos=False):
    exe = build_dropin()
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<QQQ", charges.shape[0], seeds.shape[0], length))
            f.write(np.ascontiguousarray(charges, np.float32).tobytes())
            f.write(np.ascontiguousarray(seeds, np.float32).tobytes())
        subprocess.run([exe, fin, fout, str(resolution), str(int(curvature))] + (["aos"] if aos else []),
                       check=True, timeout=120)
        raw = open(fout, "rb").read()
    st, t, g, pr, steps = struct.unpack("<idddI", raw[:32])
    rows = np.frombuffer(raw[32:], np.float32).reshape(3, length, seeds.shape[0])
    return st, t, g, pr, steps, rows


@pytest.mark.gpu
def test_dropin_program_matches_device_path(gpu):
    import torch
    from emtrace import gen, rules
    w = gen.plain_config(700, 320, 10.0, 150, seed=13)
    st, t, g, pr, steps, rows = _run_dropin(w.charges, w.seeds, w.len, 10.0, True)
    assert st == 0 and pr == 1.0 and steps == 3 and t > 0 and g > 0
    dev_rows = gpu.trace(torch.from_numpy(w.charges).cuda(), torch.from_numpy(w.seeds).cuda(),
                         w.len, rules.ref(10.0)).cpu().numpy()
    assert np.array_equal(rows, dev_rows)


@pytest.mark.gpu
def test_aos_entry_point_matches_soa(gpu):
    """emtrace::CalcField on Array<Vector3<float>> lines writes the same rows."""
    from emtrace import gen
    w = gen.plain_config(300, 192, 10.0, 90, seed=17)
    soa = _run_dropin(w.charges, w.seeds, w.len, 10.0, True)
    aos = _run_dropin(w.charges, w.seeds, w.len, 10.0, True, aos=True)
    assert soa[0] == 0 and aos[0] == 0 and aos[3] == 1.0 and aos[4] == 3
    assert np.array_equal(soa[5], aos[5])


@pytest.mark.gpu
def test_dropin_program_reports_reference_status_codes(gpu):
    from emtrace import gen
    w = gen.plain_config(10, 64, 10.0, 8, seed=1)
    assert _run_dropin(w.charges, w.seeds, w.len, 0.0, True)[0] == 2
    assert _run_dropin(w.charges, w.seeds, 1, 10.0, True)[0] == 3
