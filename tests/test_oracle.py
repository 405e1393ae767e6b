"""Pins for the oracle (CPU only).

Each oracle function is checked against something that does not share its
formula: the golden vectors written by the compiled reference, the compiled
reference itself where it is present, closed forms (a single charge), exact
invariants, and central finite differences.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from emtrace import gen, rules
from helpers import parity_report, prefix_lengths, rows64_to_soa

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden_cases():
    return sorted(p for p in glob.glob(os.path.join(GOLDEN, "ref_f*.npz")))


def _rule_of(g):
    res = float(g["resolution"])
    return rules.ref(res) if int(g["use_curvature"]) else rules.plain(res)


@pytest.mark.parametrize("path", _golden_cases(), ids=lambda p: os.path.basename(p)[:-4])
def test_restatement_matches_golden(path, orc):
    """C restatement vs the reference's own output (CPU_Implement.cpp:38-102, 207-351, 366-498)."""
    g = np.load(path)
    dt = g["rows"].dtype
    rc, rows = orc.calc_lines(g["charges"], g["seeds"], g["rows"].shape[1], _rule_of(g), dtype=dt)
    assert rc == 0
    assert np.isfinite(rows).all(), "IEEE restatement must stay finite"
    ref = g["rows"]
    first, _ = prefix_lengths(ref.astype(np.float32), g["charges"].astype(np.float32))
    err = np.abs(rows.astype(np.float64) - ref).max(axis=0)
    perr = np.array([err[: first[l], l].max() if first[l] else 0.0 for l in range(err.shape[1])])
    # row 0 is copied, and the first computed row already agrees to rounding
    assert np.array_equal(rows[:, 0], ref[:, 0])
    if dt == np.float64:
        assert perr.max() < 1e-9
    else:
        # -ffast-math reassociation in the reference build vs the literal order;
        # the drift scales with the step length 1/resolution
        step = 1.0 / float(g["resolution"])
        assert np.median(perr) < 1e-5 * step
        assert perr.max() < 5e-3 * step
        assert np.mean(rows[:, 1] == ref[:, 1]) > 0.5


def test_reference_error_codes(orc):
    """Argument errors: 1 (n == 0), 2 (resolution == 0), 3 (< 2 rows), 5 (SSE path, n % 64)."""
    g = np.load(os.path.join(GOLDEN, "ref_error_codes.npz"))
    for name, code, args in zip(g["names"], g["codes"], g["args"]):
        n, length, res, uc = int(args[0]), int(args[1]), float(args[2]), int(args[3])
        rule = rules.ref(res) if uc else rules.plain(res)
        rc, _ = orc.calc_lines(np.zeros((4, 4), np.float32), np.zeros((n, 3), np.float32),
                               length, rule)
        assert rc == code, name


@pytest.mark.skipif(oracle.ref_lib() is None, reason="compiled reference absent (oracle/_ref)")
def test_restatement_matches_live_reference(orc):
    w = gen.plain_config(300, 128, 10.0, 300, seed=11)
    for uc in (0, 1):
        rc, ref, _, _ = orc.ref_calc_lines(w.charges, w.seeds, w.len, 10.0, uc)
        rc2, rows = orc.calc_lines(w.charges, w.seeds, w.len, rules.ref(10.0) if uc else rules.plain(10.0))
        assert rc == rc2 == 0
        rep, *_ = parity_report(rows, ref, w.charges)
        assert rep["prefix_median_err"] < 1e-6
        assert rep["prefix_within_tol"] > 0.97


def test_single_charge_closed_form(orc):
    """One charge: every line is radial, x_s = c + (r0 + s*h) u (plain and ref)."""
    c = np.array([[1.25, -0.5, 2.0, 0.8]], np.float64)
    d = gen.fibonacci_dirs(64)
    seeds = c[0, :3] + 0.1 * d
    for rule in (rules.plain(10.0), rules.ref(10.0)):
        rc, rows = orc.calc_lines(c, seeds, 40, rule, dtype=np.float64)
        assert rc == 0
        s = np.arange(40)[:, None, None]
        expect = c[0, :3][None, None, :] + (0.1 + s * 0.1) * d[None, :, :]
        np.testing.assert_allclose(rows.transpose(1, 2, 0), expect, atol=1e-12)
    # a negative charge pulls lines inward by the same law
    cn = c.copy()
    cn[0, 3] = -0.8
    rc, rows = orc.calc_lines(cn, c[0, :3] + 5.0 * d, 30, rules.plain(10.0), dtype=np.float64)
    s = np.arange(30)[:, None, None]
    expect = c[0, :3][None, None, :] + (5.0 - s * 0.1) * d[None, :, :]
    np.testing.assert_allclose(rows.transpose(1, 2, 0), expect, atol=1e-12)


def test_angle_rule_single_charge(orc):
    """Straight line: chord 0 after the first step, so h = hmax from step 2 on."""
    c = np.array([[0.0, 0.0, 0.0, 1.0]], np.float64)
    d = gen.fibonacci_dirs(16)
    seeds = 0.1 * d
    rule = rules.angle()
    for dt, tol in ((np.float64, 1e-12), (np.float32, 2e-5)):
        rc, rows = orc.calc_lines(c.astype(dt), seeds.astype(dt), 20, rule, dtype=dt)
        assert rc == 0
        r = np.array([0.1, 0.2] + [0.2 + 0.5 * k for k in range(1, 19)])
        expect = r[:, None, None] * d[None, :, :]
        np.testing.assert_allclose(rows.transpose(1, 2, 0), expect, atol=tol * 10)
    r64 = orc.trace64(c, seeds, 20, rule)
    np.testing.assert_allclose(r64, r[:, None, None] * d[None], atol=1e-12)


def test_angle_rule_stops_at_a_sink(orc):
    """A line aimed at a lone sink moves straight in (h0, then hmax: chord 0),
    overshoots it, and ends at the step whose direction reverses."""
    c = np.array([[0.0, 0.0, 0.0, -1.0]], np.float64)
    d = np.array([[0.6, 0.0, 0.8]])
    seeds = 2.3 * d  # distance 2.3: 2.3 -> 2.2 -> 1.7 -> 1.2 -> 0.7 -> 0.2 -> -0.3 (past)
    rule = rules.angle()
    for dt in (np.float64, np.float32):
        rc, rows, steps = orc.calc_lines(c.astype(dt), seeds.astype(dt), 12, rule, dtype=dt,
                                         with_steps=True)
        assert rc == 0
        r = np.array([2.3, 2.2, 1.7, 1.2, 0.7, 0.2, -0.3])
        # step 7 sees the field reversed: the line ends at x_6 and repeats it
        assert int(steps[0]) == 6
        expect = np.concatenate([r, np.full(5, -0.3)])[:, None] * d
        tol = 1e-12 if dt == np.float64 else 2e-6
        np.testing.assert_allclose(rows[:, :, 0].T, expect, atol=tol)
    r64, s64 = orc.trace64(c, seeds, 12, rule, with_steps=True)
    assert int(s64[0]) == 6
    np.testing.assert_allclose(r64[:, 0, :], expect, atol=1e-12)
    # without the stop the line keeps bouncing across the sink
    rc, rows, steps = orc.calc_lines(c, seeds, 12, rules.angle(stop=False), dtype=np.float64,
                                     with_steps=True)
    assert int(steps[0]) == 11 and not np.allclose(rows[:, 7:, 0], rows[:, 6:7, 0])


def test_ref_rule_reduces_to_plain(orc):
    """k = |(E-P)xP|/|E|^3 < 2^-53 with k_e inside: (k+1) == 1 (CPU_Implement.cpp:327-331)."""
    w = gen.plain_config(64, 64, 10.0, 100, seed=2)
    _, a = orc.calc_lines(w.charges, w.seeds, w.len, rules.plain(10.0), dtype=np.float64)
    _, b = orc.calc_lines(w.charges, w.seeds, w.len, rules.ref(10.0), dtype=np.float64)
    first, _ = prefix_lengths(a.astype(np.float32), w.charges)
    err = np.abs(a - b).max(axis=0)
    perr = np.array([err[: first[l], l].max() if first[l] else 0 for l in range(64)])
    assert perr.max() < 1e-9


def test_field64_closed_form_and_linearity(orc):
    rng = np.random.default_rng(0)
    c = np.array([[0.5, -1.0, 2.0, -0.7]])
    pts = rng.uniform(-5, 5, (50, 3))
    E = orc.field64(c, pts)
    d = pts - c[0, :3]
    r = np.linalg.norm(d, axis=1, keepdims=True)
    np.testing.assert_allclose(E, -0.7 * d / r ** 3, rtol=1e-14)
    ch = np.concatenate([rng.uniform(-5, 5, (40, 3)), rng.uniform(-1, 1, (40, 1))], axis=1)
    E1 = orc.field64(ch[:20], pts)
    E2 = orc.field64(ch[20:], pts)
    np.testing.assert_allclose(orc.field64(ch, pts), E1 + E2, rtol=1e-12, atol=1e-14)
    ch2 = ch.copy()
    ch2[:, 3] *= 3.0
    np.testing.assert_allclose(orc.field64(ch2, pts), 3.0 * orc.field64(ch, pts), rtol=1e-13)
    # reference-arithmetic float path agrees with the physics form to float rounding
    Ef = orc.field_ref_f32(ch.astype(np.float32), pts.astype(np.float32), np.float32(1.0))
    np.testing.assert_allclose(Ef, orc.field64(ch.astype(np.float32).astype(np.float64),
                                               pts.astype(np.float32).astype(np.float64)),
                               rtol=2e-4, atol=1e-5)


def test_field_jvp64_matches_finite_differences(orc):
    rng = np.random.default_rng(1)
    ch = np.concatenate([rng.uniform(-3, 3, (12, 3)), rng.uniform(-1, 1, (12, 1))], axis=1)
    pts = rng.uniform(-3, 3, (20, 3))
    w = rng.normal(size=(20, 3))
    jv = orc.field_jvp64(ch, pts, w, scale=2.0)
    h = 1e-6
    fd = (orc.field64(ch, pts + h * w, 2.0) - orc.field64(ch, pts - h * w, 2.0)) / (2 * h)
    np.testing.assert_allclose(jv, fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())
    # the field-gradient tensor is symmetric and traceless away from charges
    J = np.stack([orc.field_jvp64(ch, pts, np.tile(e, (20, 1))) for e in np.eye(3)], axis=2)
    np.testing.assert_allclose(J, J.transpose(0, 2, 1), atol=1e-9 * np.abs(J).max())
    np.testing.assert_allclose(np.trace(J, axis1=1, axis2=2), 0.0, atol=1e-9 * np.abs(J).max())


@pytest.mark.parametrize("rule", [rules.plain(10.0), rules.angle()], ids=["plain", "angle"])
def test_adjoint64_matches_finite_differences(orc, rule):
    """o64_adjoint vs central differences of o64_trace (h = 1e-6, FP64)."""
    rng = np.random.default_rng(5)
    w = gen.plain_config(6, 5, 3.0, 14, seed=4)
    ch = w.charges.astype(np.float64)
    seeds = w.seeds.astype(np.float64)
    G = rng.normal(size=(w.len, w.n, 3))

    def loss(c, s):
        return float((orc.trace64(c, s, w.len, rule) * G).sum())

    gch, gseed = orc.adjoint64(ch, seeds, w.len, rule, G)
    h = 1e-6
    fd = np.zeros_like(ch)
    for i in range(ch.shape[0]):
        for k in range(4):
            cp, cm = ch.copy(), ch.copy()
            cp[i, k] += h
            cm[i, k] -= h
            fd[i, k] = (loss(cp, seeds) - loss(cm, seeds)) / (2 * h)
    np.testing.assert_allclose(gch, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    fds = np.zeros_like(seeds)
    for i in range(seeds.shape[0]):
        for k in range(3):
            sp, sm = seeds.copy(), seeds.copy()
            sp[i, k] += h
            sm[i, k] -= h
            fds[i, k] = (loss(ch, sp) - loss(ch, sm)) / (2 * h)
    np.testing.assert_allclose(gseed, fds, rtol=1e-5, atol=1e-6 * np.abs(fds).max())


@pytest.mark.parametrize("rule", [rules.plain(10.0), rules.angle()], ids=["plain", "angle"])
def test_adjoint64_path_is_the_adjoint_along_a_given_path(orc, rule):
    """Fed the FP64 trajectory itself, the path form reproduces o64_adjoint
    (up to the float rounding of the stored path)."""
    w = gen.plain_config(40, 32, 4.0, 60, seed=21)
    r64, s64 = orc.trace64(w.charges, w.seeds, w.len, rule, with_steps=True)
    G = np.random.default_rng(0).normal(size=r64.shape)
    a, sa = orc.adjoint64(w.charges, w.seeds, w.len, rule, G)
    b, sb = orc.adjoint64_path(w.charges, w.seeds, w.len, rule, G,
                               rows64_to_soa(r64).astype(np.float32), s64)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 2e-3
    assert np.linalg.norm(sa - sb) / np.linalg.norm(sa) < 2e-3


def test_restated_angle_f32_tracks_fp64(orc):
    w = gen.plain_config(200, 64, 10.0, 300, seed=9)
    rc, rows = orc.calc_lines(w.charges, w.seeds, w.len, rules.angle())
    assert rc == 0
    r64 = rows64_to_soa(orc.trace64(w.charges, w.seeds, w.len, rules.angle()))
    rep, *_ = parity_report(rows, r64, w.charges, tol=1e-3)
    # the adaptive step shifts points along the line (ill-conditioned in float),
    # the path itself stays close
    assert rep["path_median_err"] < 1e-5
    assert rep["path_within_tol"] > 0.95
    assert rep["prefix_median_err"] < 1e-4


def test_generator_is_pinned():
    """Seeded workloads are fixed by value (SplitMix64 streams, emtrace/gen.py)."""
    w0 = gen.lines_8k()
    assert w0.charges.shape == (1000, 4) and w0.seeds.shape == (8192, 3) and w0.len == 2500
    assert gen.fingerprint(w0.charges, w0.seeds) == gen.fingerprint(gen.lines_8k().charges,
                                                                     gen.lines_8k().seeds)
    assert np.all(np.abs(w0.charges[:, :3]) <= 10.0)
    mag = np.abs(w0.charges[:, 3])
    assert mag.min() >= 0.5 and mag.max() <= 1.0
    seeds_d = np.linalg.norm(w0.seeds[:, None, :] - w0.charges[None, :, :3], axis=2).min(axis=1)
    # 0.1 from the chosen source (a second charge may sit even closer)
    assert (seeds_d <= 0.1 + 2e-6).all() and (np.abs(seeds_d - 0.1) < 2e-6).mean() > 0.999
    assert gen.fingerprint(w0.charges, w0.seeds) == PINNED_CFG0
    w1 = gen.lines_64k()
    assert w1.interactions == 16384 * 65536 * 2499
    assert gen.fingerprint(w1.charges, w1.seeds) == PINNED_CFG1


PINNED_CFG0 = "b5fe415cbbb4bf85"
PINNED_CFG1 = "885a25643f9487f7"


# ---------------------------------------------------------------------------
# particle step (Particle_System.cpp:108-190): closed forms computed here
# independently of oracle64.c
KE_ = 8.987551787e9


def _kick_drift(P, a, dt):
    v = P[:, 4:7] + a * (dt / 2)
    return P[:, 0:3] + v * dt, v


def test_verlet64_coulomb_pair_with_drift(orc):
    P = np.array([[0.1, -0.2, 0.3, 1e-3, 0.5, 0.0, -0.2, 2.0],
                  [1.0, 0.5, -0.4, -2e-3, 0.0, 0.3, 0.1, 3.0]])
    a_in = np.array([[1.0, 2.0, 0.0], [0.0, -1.0, 0.5]])
    dt = 0.01
    a, _ = orc.verlet64(P, np.zeros((0, 4)), a_in, [-9] * 3, [9] * 3, dt)
    x, v = _kick_drift(P, a_in, dt)
    for i, j in ((0, 1), (1, 0)):
        r = x[i] - P[j, :3]  # the other particle at its OLD position
        r3 = np.linalg.norm(r) ** 3
        E = KE_ * P[j, 3] * r / r3
        B = 1e-7 * P[j, 3] * np.cross(P[j, 4:7], r) / r3  # Biot-Savart, point charge
        want = P[i, 3] * (E + np.cross(v[i], B)) / P[i, 7]
        assert np.allclose(a[i], want, rtol=1e-12, atol=0), (i, a[i], want)


def test_verlet64_static_charges_and_self_exclusion(orc):
    rng = np.random.default_rng(3)
    ch = np.concatenate([rng.uniform(-5, 5, (7, 3)), rng.uniform(-1, 1, (7, 1))], axis=1)
    P = np.array([[0.3, 0.2, -0.1, 0.5, 0.1, 0.2, 0.3, 1.5]])
    a, _ = orc.verlet64(P, ch, np.zeros((1, 3)), [-9] * 3, [9] * 3, 0.05)
    x, _ = _kick_drift(P, np.zeros((1, 3)), 0.05)
    r = x[0] - ch[:, :3]
    E = KE_ * (ch[:, 3:4] * r / np.linalg.norm(r, axis=1, keepdims=True) ** 3).sum(axis=0)
    assert np.allclose(a[0], P[0, 3] * E / P[0, 7], rtol=1e-12)
    a0, _ = orc.verlet64(P, np.zeros((0, 4)), np.zeros((1, 3)), [-9] * 3, [9] * 3, 0.05)
    assert np.array_equal(a0, np.zeros((1, 3)))  # a lone particle feels nothing of itself


def test_verlet64_newton_third_law_and_box(orc):
    rng = np.random.default_rng(5)
    n = 40
    P = np.zeros((n, 8))
    P[:, :3] = rng.uniform(-2, 2, (n, 3))
    P[:, 3] = rng.uniform(-1e-3, 1e-3, n)
    P[:, 7] = rng.uniform(1, 2, n)
    a, _ = orc.verlet64(P, np.zeros((0, 4)), np.zeros((n, 3)), [-9] * 3, [9] * 3, 0.0)
    F = a * P[:, 7:8]
    assert np.abs(F.sum(axis=0)).max() <= 1e-9 * np.abs(F).sum()
    # box: a particle that drifts past the upper x bound is reflected
    Q = np.array([[0.95, 0.0, 0.0, 0.0, 10.0, 0.0, 0.0, 1.0]])
    a, out = orc.verlet64(Q, np.zeros((0, 4)), np.zeros((1, 3)), [-1] * 3, [1] * 3, 0.01, True)
    assert np.allclose(out[0, :3], [2 * 1 - 1.05, 0, 0]) and np.allclose(out[0, 4:7], [-10, 0, 0])
    a, none = orc.verlet64(Q, np.zeros((0, 4)), np.zeros((1, 3)), [-1] * 3, [1] * 3, 0.01, False)
    assert none is None


# ---------------------------------------------------------------------------
# pins for the parity metrics (o64_min_dist_f32, o64_path_dist_f32) and for
# o64_trace's plain rule.  Each is checked against a formulation that does not
# share its code: numpy brute force, cross-product segment distance, geometry
# with a known answer.

def _seg_dist_cross(p, a, b):
    """Distance from p to segment ab: the end points when the foot of the
    perpendicular falls outside, else |(p - a) x (b - a)| / |b - a|."""
    ab = b - a
    L = np.linalg.norm(ab)
    da, db = np.linalg.norm(p - a), np.linalg.norm(p - b)
    if L == 0.0:
        return da
    if np.dot(p - a, ab) <= 0.0:
        return da
    if np.dot(p - b, -ab) <= 0.0:
        return db
    return np.linalg.norm(np.cross(p - a, ab)) / L


def test_min_dist_matches_brute_force(orc):
    """o64_min_dist_f32 (the "> 0.05 from any charge" filter) vs numpy."""
    rng = np.random.default_rng(12)
    ch = np.concatenate([rng.uniform(-3, 3, (23, 3)), rng.uniform(-1, 1, (23, 1))],
                        axis=1).astype(np.float32)
    rows = rng.uniform(-4, 4, (3, 11, 7)).astype(np.float32)
    got = orc.min_dist(ch, rows)
    pts = rows.transpose(1, 2, 0).astype(np.float64)            # (len, n, 3)
    want = np.sqrt(((pts[:, :, None, :] - ch[None, None, :, :3].astype(np.float64)) ** 2)
                   .sum(-1)).min(-1)
    assert got.shape == (11, 7)
    np.testing.assert_allclose(got, want, rtol=1e-6)
    # known distances: points 0.03 / 0.05 / 0.07 along the axes of an isolated
    # charge (the others 100 away)
    far = np.array([[0.5, -0.25, 1.0, 1.0], [100.0, 0, 0, 1], [0, 100.0, 0, -1]], np.float32)
    r = np.array([0.03, 0.05, 0.07])
    pts = far[0, :3] + r[:, None] * np.eye(3)                      # (3 rows, xyz)
    rows = pts.T[:, :, None].astype(np.float32)                   # (3, len=3, n=1)
    d = orc.min_dist(far, rows)[:, 0]
    np.testing.assert_allclose(d, r, rtol=2e-5)
    assert d[0] <= 0.05 + 1e-6 and d[2] > 0.05


def test_path_dist_matches_brute_force(orc):
    """With a window covering the whole line, o64_path_dist_f32 is the
    distance to the nearest segment of the reference polyline through points
    0 .. first[l] (the point where the comparison stops still bounds the last
    segment); 0 at and after first[l]."""
    rng = np.random.default_rng(13)
    length, n = 12, 5
    ref = np.cumsum(rng.normal(size=(3, length, n)), axis=1).astype(np.float32)
    test = (ref + rng.normal(scale=0.4, size=ref.shape)).astype(np.float32)
    first = np.array([12, 7, 2, 12, 5])
    got = orc.path_dist(test, ref, first, window=length)
    for l in range(n):
        for j in range(length):
            if j >= first[l]:
                assert got[j, l] == 0.0
                continue
            p = test[:, j, l].astype(np.float64)
            last = min(first[l], length - 1)
            want = min(_seg_dist_cross(p, ref[:, k, l].astype(np.float64),
                                       ref[:, k + 1, l].astype(np.float64))
                       for k in range(last))
            assert abs(got[j, l] - want) <= 1e-6 * max(1.0, want), (j, l, got[j, l], want)


def test_path_dist_geometry(orc):
    """Known answers: points on the polyline (0), at a perpendicular offset from
    a straight polyline (the offset), beyond its end (distance to the end
    point), and the window bounding the search (a circle traversed once: the
    diametrically opposite point is found only when the window reaches it)."""
    length = 40
    s = np.arange(length) * 0.1
    ref = np.zeros((3, length, 1), np.float32)
    ref[0, :, 0] = s                                          # straight along x
    first = np.array([length])
    # midpoints of the segments: on the path
    mid = ref.copy()
    mid[0, :-1, 0] = s[:-1] + 0.05
    assert orc.path_dist(mid, ref, first, 4)[:-1].max() < 1e-7
    # perpendicular offset 0.02 in y and z: distance 0.02 * sqrt(2)
    off = ref.copy()
    off[1] += 0.02
    off[2] -= 0.02
    np.testing.assert_allclose(orc.path_dist(off, ref, first, 4)[:, 0], 0.02 * np.sqrt(2), rtol=1e-5)
    # shifted along the line by 0.37 (3.7 steps): inside the window the distance
    # is 0; the last points lie beyond the end, at (s + 0.37) - s_end
    sh = ref.copy()
    sh[0] += 0.37
    d = orc.path_dist(sh, ref, first, 8)[:, 0]
    beyond = s + 0.37 > s[-1]
    assert d[~beyond].max() < 1e-6
    np.testing.assert_allclose(d[beyond], s[beyond] + 0.37 - s[-1], rtol=1e-5)
    # the same shift with a window of 2 segments: the foot lies 3.7 segments
    # ahead, outside [j-2, j+2], so the distance is to point j+2: 0.17
    d2 = orc.path_dist(sh, ref, first, 2)[:, 0]
    np.testing.assert_allclose(d2[:30], 0.17, rtol=1e-5)
    # circle of radius 1 traversed once; the test point j sits where the
    # reference is half a turn later
    ang = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    circ = np.zeros((3, 64, 1), np.float32)
    circ[0, :, 0], circ[1, :, 0] = np.cos(ang), np.sin(ang)
    opp = np.roll(circ, -32, axis=1)
    f = np.array([64])
    wide = orc.path_dist(opp, circ, f, 64)[:, 0]
    narrow = orc.path_dist(opp, circ, f, 4)[:, 0]
    assert wide.max() < 1e-6
    assert narrow.min() > 1.5      # the 4-segment window sees only points near angle j


def test_path_dist_pointwise_fallback_and_prefix_end(orc):
    """Where the window holds no segment (window 0, or a one-row line) the
    distance is pointwise, to the point of the same index.  A line compared up
    to first[l] = 1 is measured against its first segment (points 0 and 1);
    a point just past the end of the compared prefix along the path is at
    distance 0 from that segment, not at its along-line offset; a non-finite
    point at first[l] (a reference path that blew up there) is not used."""
    rng = np.random.default_rng(14)
    ref = rng.normal(size=(3, 9, 3)).astype(np.float32)
    test = rng.normal(size=(3, 9, 3)).astype(np.float32)
    pw = np.linalg.norm(test.astype(np.float64) - ref.astype(np.float64), axis=0)
    d0 = orc.path_dist(test, ref, np.array([9, 9, 9]), 0)
    np.testing.assert_allclose(d0, pw, rtol=1e-6)
    d1 = orc.path_dist(test, ref, np.array([1, 9, 1]), 3)
    for l in (0, 2):
        want = _seg_dist_cross(test[:, 0, l].astype(np.float64), ref[:, 0, l].astype(np.float64),
                               ref[:, 1, l].astype(np.float64))
        assert abs(d1[0, l] - want) <= 1e-6 * max(1.0, want)
    assert (d1[1:, [0, 2]] == 0).all()
    one = orc.path_dist(test[:, :1], ref[:, :1], np.array([1, 1, 1]), 3)
    np.testing.assert_allclose(one[0], pw[0], rtol=1e-6)
    # straight reference along x with steps 0.3; compared up to first = 5; the
    # test's point 4 lies 0.01 further along the line than the reference's
    line = np.zeros((3, 8, 1), np.float32)
    line[0, :, 0] = 0.3 * np.arange(8)
    t = line.copy()
    t[0, 4, 0] += 0.01
    assert orc.path_dist(t, line, np.array([5]), 4)[4, 0] < 1e-7
    bad = line.copy()
    bad[:, 5, 0] = 1e30                     # the reference is non-finite at first[l]
    np.testing.assert_allclose(orc.path_dist(t, bad, np.array([5]), 4)[4, 0], 0.01, rtol=1e-4)


def test_trace64_plain_closed_form(orc):
    """o64_trace, plain rule: one charge gives radial lines with
    x_s = c + (r0 + s/res) u (CPU_Implement.cpp:89-92, Vector.h:158-164),
    outward for q > 0 and inward for q < 0."""
    c = np.array([[1.25, -0.5, 2.0, 0.8]], np.float64)
    d = gen.fibonacci_dirs(48)
    for res in (10.0, 4.0):
        rows = orc.trace64(c, c[0, :3] + 0.1 * d, 60, rules.plain(res))
        s = np.arange(60)[:, None, None]
        np.testing.assert_allclose(rows, c[0, :3] + (0.1 + s / res) * d[None], atol=1e-12)
    cn = c.copy()
    cn[0, 3] = -0.8
    rows, steps = orc.trace64(cn, c[0, :3] + 5.0 * d, 30, rules.plain(10.0), with_steps=True)
    s = np.arange(30)[:, None, None]
    np.testing.assert_allclose(rows, c[0, :3] + (5.0 - s * 0.1) * d[None], atol=1e-12)
    assert (steps == 29).all()


def test_trace64_plain_matches_reference_double_rows(orc):
    """o64_trace, plain rule, vs the reference's own CalcField_CPU<double>
    rows (golden vectors written by the reference compiled from source), before
    each line's first approach within 0.05 of a charge."""
    g = np.load(os.path.join(GOLDEN, "ref_f64_plain_p64_n64_len200.npz"))
    ref = g["rows"]
    rows = rows64_to_soa(orc.trace64(g["charges"], g["seeds"], ref.shape[1],
                                     rules.plain(float(g["resolution"]))))
    first, _ = prefix_lengths(ref.astype(np.float32), g["charges"].astype(np.float32))
    err = np.abs(rows - ref).max(axis=0)
    perr = np.array([err[: first[l], l].max() if first[l] else 0.0 for l in range(ref.shape[2])])
    assert (first > 10).mean() > 0.5
    assert perr.max() < 1e-8, perr.max()


def test_trace64_jitter_pins(orc):
    """o64_trace_jitter: with no injected rounding it is o64_trace bit for bit;
    the same key repeats, another key differs; on one charge's radial lines the
    injected +-2^-24 at every step moves the path by no more than the sum of
    what was injected (the field of a point charge neither pulls a line back
    nor pushes it away from the radial through its point); and the move is
    linear in the injected size on a smooth many-charge line set."""
    w = gen.plain_config(40, 16, 4.0, 80, seed=3)
    for rule in (rules.plain(10.0), rules.angle()):
        a, sa = orc.trace64(w.charges, w.seeds, w.len, rule, with_steps=True)
        b, sb = orc.trace64_jitter(w.charges, w.seeds, w.len, rule, 0.0, 7, with_steps=True)
        assert np.array_equal(a, b) and np.array_equal(sa, sb)
        j1 = orc.trace64_jitter(w.charges, w.seeds, w.len, rule, 2.0 ** -24, 1)
        assert np.array_equal(j1, orc.trace64_jitter(w.charges, w.seeds, w.len, rule, 2.0 ** -24, 1))
        assert not np.array_equal(j1, orc.trace64_jitter(w.charges, w.seeds, w.len, rule, 2.0 ** -24, 2))
    # one charge: radial lines of 60 steps of 0.1 from r0 = 0.1
    c = np.array([[0.5, -0.25, 1.0, 1.0]])
    d = gen.fibonacci_dirs(32)
    seeds = c[0, :3] + 0.1 * d
    rel = 2.0 ** -24
    base = orc.trace64(c, seeds, 60, rules.plain(10.0))
    jit = orc.trace64_jitter(c, seeds, 60, rules.plain(10.0), rel, 3)
    dev = np.abs(jit - base).max(axis=2)                     # (len, n)
    assert dev[1:].min() > 0
    # per step: the direction by <= sqrt(3) rel (moving the point h times that)
    # and each coordinate by rel |x_k|; the largest |x_k| is <= |c_k| + r
    reach = np.abs(c[0, :3]).max() + 0.1 + 0.1 * np.arange(60)
    bound = np.cumsum(np.sqrt(3) * rel * 0.1 + rel * reach)
    assert (dev <= bound[:, None] * 1.01).all()
    # linear in the injected size (the perturbation stays in the linear regime)
    base = orc.trace64(w.charges, w.seeds, w.len, rules.plain(10.0))
    m1 = np.abs(orc.trace64_jitter(w.charges, w.seeds, w.len, rules.plain(10.0), 1e-9, 5) - base)
    m2 = np.abs(orc.trace64_jitter(w.charges, w.seeds, w.len, rules.plain(10.0), 2e-9, 5) - base)
    first, _ = prefix_lengths(rows64_to_soa(base).astype(np.float32), w.charges)
    for l in range(w.n):
        f = first[l]
        if f > 5:
            r = m2[:f, l].max() / m1[:f, l].max()
            assert abs(r - 2.0) < 0.05, (l, r)
