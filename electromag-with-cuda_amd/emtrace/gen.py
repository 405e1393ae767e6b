"""Seeded synthetic workloads for the field-line tracer.

The reference initialises its problem with libc ``rand()`` under ``srand(1)``
(ElectroMag/src/Electromag utils.h:29-111).  The workloads here are defined
independently of any libc: a counter-based SplitMix64 stream per purpose, so
that the same seed gives the same charges and seeds in numpy, on any host.

Charges are float32 records ``(x, y, z, q)`` -- the layout of
``electro::pointCharge<float>`` (include/Electrostatics.h:30-35).  Seeds are the
row-0 points of the step-major line buffer (``x/y/z[step*n + line]``,
ElectroMag/src/CPU_Implement.cpp:80,92).

Workloads (sizes are those of the benchmark configurations):

* ``plain_config(p, n, half_side)``  -- positions U[-a, a]^3, magnitudes
  U[0.5, 1.0] with a random sign; every line seeded 0.1 from a uniformly chosen
  positive charge, direction on a Fibonacci sphere over the line index.
* ``batched_configs(...)``          -- many independent configurations with
  log-uniform charge and line counts, cube side 20*(p/1000)^(1/3) for p
  charges (the charge density of the 1,000-charge plain workload).
* ``perturbed(charges, sigma, scale_lo, scale_hi)`` -- target charges for the
  inverse problem: positions + N(0, sigma^2), magnitudes * U[lo, hi].
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# stream identifiers
S_POS, S_MAG, S_SIGN, S_PICK, S_CFG_P, S_CFG_N, S_PERT_POS, S_PERT_Q = range(1, 9)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def raw_stream(seed: int, stream: int, count: int) -> np.ndarray:
    """``count`` uint64 values of SplitMix64 stream ``stream`` under ``seed``."""
    key = _mix64(np.array([(seed & 0xFFFFFFFF) << 32 | (stream & 0xFFFFFFFF)],
                          dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        ctr = key + (np.arange(1, count + 1, dtype=np.uint64) * _GOLDEN)
    return _mix64(ctr)


def uniform(seed: int, stream: int, count: int) -> np.ndarray:
    """Float64 values in [0, 1) carrying 24 random bits (exact in float32)."""
    return (raw_stream(seed, stream, count) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24


def normal(seed: int, stream: int, count: int) -> np.ndarray:
    """Box-Muller normals from a 53-bit uniform stream."""
    m = (count + 1) // 2
    r = raw_stream(seed, stream, 2 * m)
    u1 = ((r[:m] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (r[m:] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)])
    return z[:count]


def make_charges(p: int, half_side: float, seed: int = 0) -> np.ndarray:
    """(p, 4) float32: positions U[-a, a]^3, magnitudes +-U[0.5, 1.0]."""
    pos = uniform(seed, S_POS, 3 * p).reshape(3, p).T
    mag = 0.5 + 0.5 * uniform(seed, S_MAG, p)
    sign = np.where(raw_stream(seed, S_SIGN, p) >> np.uint64(63), -1.0, 1.0)
    q = np.empty((p, 4), dtype=np.float32)
    q[:, :3] = (-half_side + 2.0 * half_side * pos).astype(np.float32)
    q[:, 3] = (sign * mag).astype(np.float32)
    if p and not (q[:, 3] > 0).any():  # guarantee one source to seed from
        q[0, 3] = abs(q[0, 3])
    return q


def fibonacci_dirs(n: int) -> np.ndarray:
    """(n, 3) float64 unit vectors: z = 1-(2k+1)/n, phi = k*pi*(3-sqrt 5)."""
    k = np.arange(n, dtype=np.float64)
    zc = 1.0 - (2.0 * k + 1.0) / max(n, 1)
    r = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
    phi = k * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), zc], axis=1)


def make_seeds(charges: np.ndarray, n: int, radius: float = 0.1, seed: int = 0) -> np.ndarray:
    """(n, 3) float32 seeds, each ``radius`` from a uniformly chosen positive charge."""
    pos_idx = np.nonzero(charges[:, 3] > 0)[0]
    pick = (uniform(seed, S_PICK, n) * len(pos_idx)).astype(np.int64)
    src = charges[pos_idx[pick], :3].astype(np.float64)
    return (src + radius * fibonacci_dirs(n)).astype(np.float32)


@dataclass
class Workload:
    charges: np.ndarray  # (p, 4) float32
    seeds: np.ndarray    # (n, 3) float32
    len: int             # rows per line (row 0 = seeds)

    @property
    def p(self) -> int:
        return int(self.charges.shape[0])

    @property
    def n(self) -> int:
        return int(self.seeds.shape[0])

    @property
    def interactions(self) -> int:
        return self.p * self.n * (self.len - 1)


def plain_config(p: int, n: int, half_side: float, length: int, seed: int = 0) -> Workload:
    ch = make_charges(p, half_side, seed)
    return Workload(ch, make_seeds(ch, n, seed=seed), length)


def lines_8k(seed: int = 0) -> Workload:
    """1,000 charges in [-10,10]^3, 8,192 lines, 2,500 rows."""
    return plain_config(1000, 8192, 10.0, 2500, seed)


def lines_64k(seed: int = 0) -> Workload:
    """16,384 charges in [-50,50]^3, 65,536 lines, 2,500 rows."""
    return plain_config(16384, 65536, 50.0, 2500, seed)


def adjoint_2k(seed: int = 0) -> Workload:
    """256 charges in [-5,5]^3, 2,048 lines, 500 rows (differentiation)."""
    return plain_config(256, 2048, 5.0, 500, seed)


def log_uniform_int(seed: int, stream: int, count: int, lo: int, hi: int) -> np.ndarray:
    u = uniform(seed, stream, count)
    return np.rint(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)


def batched_configs(count: int, p_range, n_range, length: int, seed: int = 0):
    """``count`` independent workloads; config i uses seed ``seed*1_000_003 + i + 1``.

    Configuration i's charges lie in a cube of side 20*(p_i/1000)^(1/3), p_i
    being its charge count: the batched workload's description gives the side
    as 20*(n/1000)^(1/3), and n is read as the charge count, so that every
    configuration keeps the charge density of the 1,000-charge, side-20 plain
    workload (1,000 charges in [-10, 10]^3) whatever its line count."""
    ps = log_uniform_int(seed, S_CFG_P, count, *p_range)
    ns = log_uniform_int(seed, S_CFG_N, count, *n_range)
    out = []
    for i in range(count):
        p, n = int(ps[i]), int(ns[i])
        sub = seed * 1_000_003 + i + 1
        half = 10.0 * (p / 1000.0) ** (1.0 / 3.0)
        out.append(plain_config(p, n, half, length, sub))
    return out


def batch_512(seed: int = 0):
    """512 configurations, p in [64, 8192], n in [256, 4096], 1,000 rows."""
    return batched_configs(512, (64, 8192), (256, 4096), 1000, seed)


def fit_64(seed: int = 0):
    """64 configurations, p in [128, 1024], n in [512, 2048], 800 rows."""
    return batched_configs(64, (128, 1024), (512, 2048), 800, seed)


def perturbed(charges: np.ndarray, sigma: float, lo: float = 0.95, hi: float = 1.05,
              seed: int = 0) -> np.ndarray:
    p = charges.shape[0]
    out = charges.astype(np.float64).copy()
    out[:, :3] += sigma * normal(seed, S_PERT_POS, 3 * p).reshape(p, 3)
    out[:, 3] *= lo + (hi - lo) * uniform(seed, S_PERT_Q, p)
    return out.astype(np.float32)


def fingerprint(*arrays: np.ndarray) -> str:
    """Short content hash used to pin the generator's output in tests."""
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]
