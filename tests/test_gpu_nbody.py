"""Particle step on the GPU (em_verlet_step_*) against the FP64 oracle.

* new accelerations within FP32 tolerance of the FP64 step: per particle
  |a_gpu - a64| <= 1e-5 |a64| on >= 99 % of particles, median <= 1e-6
  (relative to the sum of the magnitudes of the terms everywhere);
* single- and multi-tile static and dynamic sources, partial workgroups, no
  static charges, a lone particle;
* the written-back state (positions, velocities, box reflection) matches;
  without write-back the particles are left unchanged, as in the reference;
* host and device forms give identical bits; status codes.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _system(n, p, seed, box=5.0):
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 8), np.float32)
    P[:, :3] = rng.uniform(-box, box, (n, 3))
    P[:, 3] = rng.uniform(-1e-3, 1e-3, n)
    P[:, 4:7] = rng.normal(0, 1.0, (n, 3))
    P[:, 7] = rng.uniform(0.5, 2.0, n)
    ch = np.zeros((p, 4), np.float32)
    ch[:, :3] = rng.uniform(-box, box, (p, 3))
    ch[:, 3] = rng.uniform(-1e-3, 1e-3, p)
    a = rng.normal(0, 10.0, (n, 3)).astype(np.float32)
    return P, ch, a


def _check_acc(a_gpu, a64, tag):
    err = np.linalg.norm(a_gpu.astype(np.float64) - a64, axis=1)
    rel = err / np.maximum(np.linalg.norm(a64, axis=1), 1e-300)
    print(f"[nbody {tag}] rel err median {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e} "
          f"max {rel.max():.2e}", flush=True)
    assert np.isfinite(a_gpu).all()
    assert np.median(rel) <= 1e-6
    assert np.quantile(rel, 0.99) <= 1e-5


CASES = [
    # (name, n, p)
    ("small", 300, 200),
    ("ragged_multitile", 5000, 5000),
    ("no_static", 1000, 0),
    ("lone_particle", 1, 64),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_accelerations_vs_fp64(gpu, orc, case):
    name, n, p = case
    P, ch, a = _system(n, p, seed=n + p)
    dt = 1e-3
    a64, out64 = orc.verlet64(P, ch, a, [-6] * 3, [6] * 3, dt, write_back=True)
    acc = a.copy()
    Pg = P.copy()
    perf = gpu.verlet_step(Pg, ch, acc, [-6] * 3, [6] * 3, dt, write_back=True)
    assert perf["progress"] == 1.0 and perf["interactions"] == n * (p + n - 1)
    _check_acc(acc, a64, name)
    # advanced state: positions to FP32 rounding, velocities through a dt/2 kick
    assert np.allclose(Pg[:, :3], out64[:, :3], rtol=1e-6, atol=1e-6)
    assert np.array_equal(Pg[:, 3], P[:, 3]) and np.array_equal(Pg[:, 7], P[:, 7])
    vtol = 1e-5 * np.abs(a64).max() * dt + 1e-5
    assert np.abs(Pg[:, 4:7] - out64[:, 4:7]).max() <= vtol


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_lanes_per_particle_vs_fp64(gpu, orc, monkeypatch, lanes):
    """The sum over sources split over 1-8 lanes per particle (the library
    picks the count from the grid size; EMTRACE_VERLET_LANES overrides it):
    every split meets the FP64 bar, single- and multi-tile."""
    n, p = 3000, 5000
    P, ch, a = _system(n, p, seed=77)
    dt = 1e-3
    a64, _ = orc.verlet64(P, ch, a, [-6] * 3, [6] * 3, dt, write_back=False)
    monkeypatch.setenv("EMTRACE_VERLET_LANES", str(lanes))
    acc = a.copy()
    gpu.verlet_step(P.copy(), ch, acc, [-6] * 3, [6] * 3, dt, write_back=False)
    _check_acc(acc, a64, f"lanes {lanes}")


def test_bad_lanes_override_is_rejected(gpu, monkeypatch):
    from emtrace._lib import EmError
    P, ch, a = _system(64, 10, seed=5)
    monkeypatch.setenv("EMTRACE_VERLET_LANES", "3")
    with pytest.raises(EmError):
        gpu.verlet_step(P.copy(), ch, a.copy(), [-6] * 3, [6] * 3, 1e-3, write_back=False)


def test_reference_semantics_and_device_form(gpu, orc):
    P, ch, a = _system(700, 300, seed=9)
    acc_h = a.copy()
    Ph = P.copy()
    gpu.verlet_step(Ph, ch, acc_h, [-6] * 3, [6] * 3, 2e-3, write_back=False)
    assert np.array_equal(Ph, P)  # the reference only writes the accelerations
    dP = torch.from_numpy(P).cuda()
    dA = torch.from_numpy(a).cuda()
    out = torch.empty_like(dP)
    gpu.verlet_step_dev(dP, torch.from_numpy(ch).cuda(), dA, [-6] * 3, [6] * 3, 2e-3, out=out)
    torch.cuda.synchronize()
    assert np.array_equal(dA.cpu().numpy(), acc_h)
    acc2 = a.copy()
    Pw = P.copy()
    gpu.verlet_step(Pw, ch, acc2, [-6] * 3, [6] * 3, 2e-3, write_back=True)
    assert np.array_equal(out.cpu().numpy(), Pw)


def test_box_reflection(gpu):
    P = np.zeros((2, 8), np.float32)
    P[0] = [0.95, -0.95, 0.0, 0.0, 10.0, -10.0, 0.0, 1.0]  # leaves through +x and -y
    P[1] = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]
    acc = np.zeros((2, 3), np.float32)
    gpu.verlet_step(P, np.zeros((0, 4), np.float32), acc, [-1] * 3, [1] * 3, 0.01, write_back=True)
    assert np.allclose(P[0, :3], [0.95, -0.95, 0.0], atol=1e-6)
    assert np.allclose(P[0, 4:7], [-10.0, 10.0, 0.0])


def test_status_codes(gpu):
    import ctypes
    from emtrace import EmError
    from emtrace._lib import lib
    L = lib()
    b = (ctypes.c_float * 3)(0, 0, 0)
    assert L.em_verlet_step_f32(None, 0, None, 0, None, b, b, ctypes.c_float(0.1), 0, None) == 1
    dP = torch.zeros((4, 8), device="cuda")
    with pytest.raises(EmError):
        gpu.verlet_step_dev(dP, torch.zeros((0, 4), device="cuda"), torch.zeros((4, 3), device="cuda"),
                            [-1] * 3, [1] * 3, 0.1, out=dP)
