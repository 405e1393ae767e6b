"""em_trace_dev_f32_noalloc: the allocation-free form of em_trace_dev_f32.

Called through the C ABI, it must give the same rows and step counts as
em_trace_dev_f32, bit for bit (the same kernels with the same launch shape),
on two calls in a row where the second is smaller -- the fixed work array then
holds the first call's contents -- and on the paths that need no work memory.
Problems whose work memory exceeds the fixed array are refused before any
launch, leaving the rows untouched.
"""
import ctypes

import numpy as np
import pytest
import torch

import emtrace
from emtrace import _lib, gen, rules

pytestmark = pytest.mark.gpu

LIMIT = 8 << 20  # EM_NOALLOC_SCRATCH_BYTES


def _pair(gpu, w, rule, with_steps, pitch=None):
    """Trace the same problem through em_trace_dev_f32 and the allocation-free form."""
    c = torch.from_numpy(w.charges).cuda()
    s = torch.from_numpy(w.seeds).cuda()
    out = []
    for noalloc in (False, True):
        rows = gpu.seed_rows(s, w.len, pitch)
        steps = torch.full((w.n,), -1, dtype=torch.int32, device="cuda") if with_steps else None
        gpu.trace_into(c, rows, w.n, rule, steps=steps, noalloc=noalloc)
        torch.cuda.synchronize()
        out.append((rows[:, :, :w.n].cpu().numpy(), steps.cpu().numpy() if with_steps else None))
    return out


def _same(a, b):
    (ra, sa), (rb, sb) = a, b
    assert np.isfinite(ra).all()
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    if sa is not None:
        assert np.array_equal(sa, sb)


def test_noalloc_equals_trace_dev_on_two_calls(gpu):
    """Line ends (angle rule with stop, >= 4,096 charges: the register-resident
    trace over the fixed array), a larger call then a smaller one."""
    rule = rules.angle()
    big = gen.plain_config(6000, 4096, 20.0, 400, seed=31)
    print("[noalloc] call 1: p 6000, n 4096, len 400, steps kept", flush=True)
    a, b = _pair(gpu, big, rule, True)
    _same(a, b)
    assert (a[1] >= 1).all() and (a[1] > 192).any()  # lines that outlast the settling pass
    small = gen.plain_config(4200, 1500, 15.0, 260, seed=32)
    print("[noalloc] call 2: p 4200, n 1500, len 260, steps in the array, pitch 1536",
          flush=True)
    a, b = _pair(gpu, small, rule, False, pitch=1536)
    _same(a, b)


def test_noalloc_lane_group_trace_on_two_calls(gpu):
    """Above 16,384 charges the lane-group trace (set-aside list and padded
    charge copy in the fixed array), a larger call then a smaller one."""
    rule = rules.angle()
    big = gen.plain_config(16500, 2048, 40.0, 300, seed=35)
    print("[noalloc] call 1: p 16500, n 2048, len 300, steps kept", flush=True)
    a, b = _pair(gpu, big, rule, True)
    _same(a, b)
    small = gen.plain_config(16400, 700, 40.0, 220, seed=36)
    print("[noalloc] call 2: p 16400, n 700, len 220, steps in the array", flush=True)
    a, b = _pair(gpu, small, rule, False)
    _same(a, b)


def test_noalloc_paths_without_work_memory(gpu):
    """Plain rule and line ends below 4,096 charges: no work memory, same rows."""
    w = gen.plain_config(700, 900, 10.0, 120, seed=33)
    for rule in (rules.plain(10.0), rules.angle()):
        print(f"[noalloc] {rules.NAMES[rule.rule]} p 700 n 900 len 120", flush=True)
        a, b = _pair(gpu, w, rule, True)
        _same(a, b)


def _refused(gpu, p, n, with_steps):
    length = 2
    w = gen.plain_config(p, 64, 20.0, length, seed=34)
    c = torch.from_numpy(w.charges).cuda()
    seeds = torch.from_numpy(np.resize(w.seeds, (n, 3)).astype(np.float32)).cuda()
    rows = gpu.seed_rows(seeds, length)
    rows[:, 1:] = 7.0
    steps = torch.empty(n, dtype=torch.int32, device="cuda") if with_steps else None
    with pytest.raises(_lib.EmError) as e:
        gpu.trace_into(c, rows, n, rules.angle(), steps=steps, noalloc=True)
    assert e.value.code == _lib.EM_ERR_INVALID
    assert "fixed" in str(e.value)
    torch.cuda.synchronize()
    assert bool((rows[:, 1:] == 7.0).all())
    # the same problem through the allocating form is accepted
    gpu.trace_into(c, rows, n, rules.angle(), steps=steps)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(rows).all())


def test_noalloc_refuses_beyond_the_fixed_array(gpu):
    """Problems whose work memory exceeds the fixed 8 MiB are refused with
    EM_ERR_INVALID before any launch (rows 1.. keep their sentinel): the
    lane-group trace (16,500 charges) needs 48 B of set-aside record per line,
    so 200,000 lines are too many; the register-resident trace (4,096
    charges) needs a fixed 416 KiB of lists on 256 CUs plus 4 B per line of
    step counts when the caller keeps none, so 2,200,000 lines without a
    steps buffer are too many, and with one they fit."""
    assert 256 + 48 * 200_000 > LIMIT
    print("[noalloc] refuse: p 16500 n 200000", flush=True)
    _refused(gpu, 16500, 200_000, True)
    assert 4 * 2_200_000 > LIMIT
    print("[noalloc] refuse: p 4096 n 2200000, no steps buffer", flush=True)
    _refused(gpu, 4096, 2_200_000, False)
