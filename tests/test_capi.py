"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/emtrace.h, its structs match the ctypes mirror, and the
argument checks return the reference's status codes before touching a device."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import emtrace
from emtrace import _lib

HEADER = _lib.HEADER_PATH


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(em_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = emtrace.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "emtrace.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(em_step), sizeof(em_launch), sizeof(em_perf),
         sizeof(em_batch_cfg), offsetof(em_perf, progress), offsetof(em_batch_cfg, p));
  return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    want = [ctypes.sizeof(_lib.em_step), ctypes.sizeof(_lib.em_launch), ctypes.sizeof(_lib.em_perf),
            ctypes.sizeof(_lib.em_batch_cfg), _lib.em_perf.progress.offset, _lib.em_batch_cfg.p.offset]
    assert got == want


def _arrays(n, length):
    return [np.zeros(n * length, np.float32) for _ in range(3)]


def test_argument_errors_without_device():
    """Status codes of CPU_Implement.cpp:43-52 come before any device work."""
    L = emtrace.lib()
    ch = np.zeros((4, 4), np.float32)
    x, y, z = _arrays(64, 10)
    P = lambda a: a.ctypes.data  # noqa: E731
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 0, 10, 10.0, 1, None) == 1
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, 0.0, 1, None) == 2
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, 0.0, 0, None) == 2
    assert L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 64, 1, 10.0, 1, None) == 3
    assert b"len" in L.em_last_error()
    st = _lib.em_step(7, 10.0, 0.1, 0.005, 0.5, 0.05)
    assert L.em_trace_host_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, ctypes.byref(st), None, None, None) == 16
    st = _lib.em_step(2, 10.0, 0.1, 0.6, 0.5, 0.05)  # hmin > hmax
    assert L.em_trace_host_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, ctypes.byref(st), None, None, None) == 16
    st = _lib.em_step(0, 10.0, 0.1, 0.005, 0.5, 0.05, 1)  # stop is an angle-rule option
    assert L.em_trace_host_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, ctypes.byref(st), None, None, None) == 16
    assert b"angle" in L.em_last_error()
    assert L.em_trace_host_f32(None, 4, P(x), P(y), P(z), 64, 10, ctypes.byref(
        _lib.em_step(2, 10.0, 0.1, 0.005, 0.5, 0.05, 1)), None, None, None) in (16, 19)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU driver is present")
def test_no_device_is_reported():
    L = emtrace.lib()
    assert L.em_device_count() == 0
    ch = np.zeros((4, 4), np.float32)
    x, y, z = _arrays(64, 10)
    P = lambda a: a.ctypes.data  # noqa: E731
    rc = L.em_calc_lines_f32(P(ch), 4, P(x), P(y), P(z), 64, 10, 10.0, 1, None)
    assert rc == 19, L.em_last_error()
    with pytest.raises(emtrace.EmError):
        emtrace.trace_host(ch, np.zeros((64, 3), np.float32), 10, emtrace.rules.angle())


def test_python_layer_refuses_without_library(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libemtrace.so")
    with pytest.raises(emtrace.LibraryMissing):
        _lib.lib()
