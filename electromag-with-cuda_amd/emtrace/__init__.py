"""emtrace -- MI355X (gfx950) electrostatic field-line tracer.

Host-side mirror of the reference's calc-lines interface over the C ABI of
``lib/libemtrace.so`` (include/emtrace.h).  The compute runs only in the HIP
library; importing this package never falls back to a CPU path.
"""
from . import gen, legacy, rules
from ._lib import EmError, LibraryMissing, lib, LIB_PATH
from .api import (BatchLayout, BatchPlan, calc_lines, calc_lines64, ckpt_size, field, make_launch,
                  make_step, pack_batch, seed_rows, trace, trace64, trace_host, trace_host64,
                  trace_into, verlet_step, verlet_step_dev)
from .rules import ANGLE, PLAIN, REF, StepRule
from .autograd import adjoint, batch_trace, trace_lines

__all__ = [
    "gen", "legacy", "rules", "EmError", "LibraryMissing", "lib", "LIB_PATH", "BatchLayout", "BatchPlan",
    "calc_lines", "calc_lines64", "ckpt_size", "field", "make_launch", "make_step", "pack_batch", "seed_rows",
    "trace", "trace64", "trace_host", "trace_host64", "trace_into", "verlet_step", "verlet_step_dev", "ANGLE", "PLAIN", "REF", "StepRule", "adjoint",
    "batch_trace", "trace_lines",
]
