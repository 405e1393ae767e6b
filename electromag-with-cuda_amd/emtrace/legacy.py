"""The reference application's own problems (ElectroMag presets).

ElectroMag sizes its run from a preset (ElectroMag/src/ElectroMag.cpp:33-58:
lines nx x ny x nz, static charges p, rows len) and fills it with
InitializePointChargeArray / InitializeFieldLineArray
(ElectroMag/src/Electromag utils.h:29-111): charges from libc rand() after
srand(1), line seeds on a grid spaced 10 apart.  The generator runs in the
library (em_legacy_problem_f32, a built-in copy of glibc's rand()), so these
problems are the reference's bit for bit on any host.
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np

from ._lib import call
from .gen import Workload

# name: (nx, ny, nz, p, len) -- ElectroMag.cpp:43-58 (pDynamic is always 0)
PRESETS: Dict[str, Tuple[int, int, int, int, int]] = {
    "default": (128, 128, 1, 1024, 2500),
    "enhanced": (256, 112, 1, 2048, 5000),
    "extreme": (256, 256, 1, 2048, 5000),
    "insane": (512, 512, 1, 2048, 5000),
    "fuckingInsane": (1024, 1024, 1, 5120, 10000),
    "cpu": (64, 64, 1, 1000, 1000),
    "micro": (16, 16, 1, 1000, 1000),
    "bogo": (16, 16, 1, 50, 500),
}


def rand(seed: int, count: int) -> np.ndarray:
    """The first ``count`` values of rand() after srand(seed) (glibc)."""
    out = np.empty(count, np.int32)
    call("em_legacy_rand", seed & 0xFFFFFFFF, out.ctypes.data, count)
    return out


def problem(width: int, height: int, depth: int, p: int, length: int, charge_seed: int = 1,
            random_lines: bool = False, line_seed: int = 1) -> Workload:
    """Charges and seeds as ElectroMag initialises them (float32)."""
    n = width * height * depth
    charges = np.empty((p, 4), np.float32)
    xyz = np.empty((3, n), np.float32)
    call("em_legacy_problem_f32", charges.ctypes.data, p, charge_seed & 0xFFFFFFFF,
         xyz[0].ctypes.data, xyz[1].ctypes.data, xyz[2].ctypes.data, width, height, depth,
         int(random_lines), line_seed & 0xFFFFFFFF)
    return Workload(charges, np.ascontiguousarray(xyz.T), length)


def preset(name: str = "default") -> Workload:
    """A preset's problem with the reference's fixed seed (srand(1), grid seeds)."""
    nx, ny, nz, p, length = PRESETS[name]
    return problem(nx, ny, nz, p, length)
