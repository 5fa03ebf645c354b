"""Pins for the CPU oracle (no GPU).  The reference cannot be built (CUDA
clusters, and two compile errors at src/sim.cuh:147,173), so every function of
the oracle is pinned by what the mathematics fixes: closed forms evaluated in
40-digit arithmetic, conservation laws, identities and convergence orders.
Each check is chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails it."""
import math

import mpmath as mp
import numpy as np
import pytest

import oracle as O

mp.mp.dps = 40


# --- Stumpff functions --------------------------------------------------------
def _stumpff_mp(z):
    z = mp.mpf(z)
    if z > 0:
        s = mp.sqrt(z)
        return [mp.cos(s), mp.sin(s) / s, (1 - mp.cos(s)) / z, (s - mp.sin(s)) / (s * z)]
    if z < 0:
        s = mp.sqrt(-z)
        return [mp.cosh(s), mp.sinh(s) / s, (mp.cosh(s) - 1) / (-z), (mp.sinh(s) - s) / (s * (-z))]
    return [mp.mpf(1), mp.mpf(1), mp.mpf(1) / 2, mp.mpf(1) / 6]


@pytest.mark.parametrize("z", [0.0, 1e-12, -3e-7, 0.05, -0.09, 0.1, 0.11, 0.7, -2.5, 9.0, 30.0, -40.0, 100.0])
def test_stumpff_closed_forms(z):
    c = O.stumpff(z)
    ref = _stumpff_mp(z)
    for k in range(4):
        scale = max(abs(float(ref[k])), 1e-3)
        # quartering/doubling amplifies rounding ~ 4^n; n <= 5 for |z| <= 100
        tol = 1e-15 if abs(z) <= 0.1 else 2e-13
        assert abs(c[k] - float(ref[k])) <= tol * scale, (k, c[k], ref[k])


# --- Kepler drift ---------------------------------------------------------------
def _drift_mp(x, v, mu, dt):
    """Universal-variable drift in 40-digit arithmetic with closed-form
    Stumpff functions and mpmath's root finder (independent of the oracle's
    series and Newton)."""
    x = [mp.mpf(a) for a in x]
    v = [mp.mpf(a) for a in v]
    mu = mp.mpf(mu)
    dt = mp.mpf(dt)
    r0 = mp.sqrt(sum(a * a for a in x))
    eta0 = sum(a * b for a, b in zip(x, v))
    beta = 2 * mu / r0 - sum(a * a for a in v)
    zeta0 = mu - beta * r0

    def G(s):
        c = _stumpff_mp(beta * s * s)
        return c[0], s * c[1], s * s * c[2], s ** 3 * c[3]

    def F(s):
        _, g1, g2, g3 = G(s)
        return r0 * s + eta0 * g2 + zeta0 * g3 - dt

    s = mp.findroot(F, dt / r0)
    g0, g1, g2, g3 = G(s)
    r = r0 * g0 + eta0 * g1 + mu * g2
    f = 1 - mu * g2 / r0
    g = dt - mu * g3
    fd = -mu * g1 / (r0 * r)
    gd = 1 - mu * g2 / r
    assert abs(f * gd - fd * g - 1) < mp.mpf(10) ** -30  # Lagrange identity in the reference solve
    return [float(f * a + g * b) for a, b in zip(x, v)], [float(fd * a + gd * b) for a, b in zip(x, v)]


def _state_from_elements(a, e, inc, node, argp, nu, mu=1.0):
    p = a * (1 - e * e)
    r = p / (1 + e * math.cos(nu))
    xp, yp = r * math.cos(nu), r * math.sin(nu)
    h = math.sqrt(mu * abs(p))
    vxp, vyp = -mu / h * math.sin(nu), mu / h * (e + math.cos(nu))
    ci, si, co, so, cw, sw = map(lambda f: f, (math.cos(inc), math.sin(inc), math.cos(node), math.sin(node),
                                                 math.cos(argp), math.sin(argp)))
    P = np.array([cw * co - sw * so * ci, cw * so + sw * co * ci, sw * si])
    Q = np.array([-sw * co - cw * so * ci, -sw * so + cw * co * ci, cw * si])
    return xp * P + yp * Q, vxp * P + vyp * Q


CASES = [
    # a, e, i, node, argp, true anomaly, dt
    (600.0, 0.93, 0.5, 1.0, 2.0, 0.3, 0.5),      # distant TP near perihelion
    (150.0, 0.7, 0.1, 4.0, 1.0, 3.0, 0.5),       # TP near aphelion
    (5.2, 0.048, 0.02, 1.7, 4.8, 1.0, 0.5),      # Jupiter-like, n dt ~ 0.26
    (1.0, 0.017, 0.0, 0.0, 1.0, 2.0, 1e-3),      # Earth-like
    (1.0, 0.0, 0.3, 0.2, 0.0, 0.0, 2.0),         # circular, a third of an orbit
    (-20.0, 1.5, 0.4, 0.1, 0.2, 0.5, 3.0),       # hyperbolic
    (3.0, 0.999, 0.2, 0.0, 0.0, 2.5, 0.7),       # near-parabolic ellipse
]


@pytest.mark.parametrize("guess", [0, 1])
@pytest.mark.parametrize("case", CASES)
def test_kepler_drift_vs_40_digit_solution(case, guess):
    a, e, inc, node, argp, nu, dt = case
    x, v = _state_from_elements(a, e, inc, node, argp, nu)
    xo, vo, it, _ = O.kepler_drift(x, v, 1.0, dt, guess)
    assert it > 0, "Newton hit its cap"
    xr, vr = _drift_mp(x, v, 1.0, dt)
    sx = np.linalg.norm(xr)
    sv = np.linalg.norm(vr)
    assert np.max(np.abs(xo - xr)) <= 1e-13 * sx
    assert np.max(np.abs(vo - vr)) <= 1e-13 * sv


def test_kepler_drift_conserves_energy_and_angular_momentum():
    rng = np.random.default_rng(7)
    for _ in range(200):
        a = rng.uniform(2, 1000)
        e = rng.uniform(0, 0.97)
        x, v = _state_from_elements(a, e, rng.uniform(0, 1), rng.uniform(0, 6), rng.uniform(0, 6),
                                    rng.uniform(0, 6))
        dt = rng.uniform(0.01, 5.0)
        xo, vo, it, _ = O.kepler_drift(x, v, 1.0, dt, 1)
        E0 = 0.5 * v @ v - 1 / np.linalg.norm(x)
        E1 = 0.5 * vo @ vo - 1 / np.linalg.norm(xo)
        h0 = np.cross(x, v)
        h1 = np.cross(xo, vo)
        assert abs(E1 - E0) <= 1e-13 * abs(E0)
        assert np.linalg.norm(h1 - h0) <= 1e-13 * np.linalg.norm(h0)


def test_kepler_drift_composes_and_reverses():
    x, v = _state_from_elements(40.0, 0.3, 0.2, 1.0, 2.0, 0.4)
    x1, v1, _, _ = O.kepler_drift(x, v, 1.0, 0.3, 1)
    x2, v2, _, _ = O.kepler_drift(x1, v1, 1.0, 0.2, 1)
    x3, v3, _, _ = O.kepler_drift(x, v, 1.0, 0.5, 1)
    assert np.allclose(x2, x3, rtol=0, atol=1e-13 * np.linalg.norm(x3))
    xb, vb, _, _ = O.kepler_drift(x3, v3, 1.0, -0.5, 1)
    assert np.allclose(xb, x, rtol=0, atol=1e-13 * np.linalg.norm(x))
    assert np.allclose(vb, v, rtol=0, atol=1e-13 * np.linalg.norm(v))


def test_circular_orbit_rotates_by_n_dt():
    r, dt = 9.0, 1.7
    n = math.sqrt(1 / r ** 3)
    x, v, _, _ = O.kepler_drift([r, 0, 0], [0, math.sqrt(1 / r), 0], 1.0, dt, 1)
    assert abs(x[0] - r * math.cos(n * dt)) < 1e-14 * r
    assert abs(x[1] - r * math.sin(n * dt)) < 1e-14 * r


def test_series_guess_needs_fewer_newton_iterations_than_crude():
    x, v = _state_from_elements(300.0, 0.85, 0.3, 1.0, 2.0, 0.2)
    _, _, it_crude, s0 = O.kepler_drift(x, v, 1.0, 0.5, 0)
    _, _, it_series, s1 = O.kepler_drift(x, v, 1.0, 0.5, 1)
    assert it_series < it_crude
    assert abs(s0 - s1) <= 1e-15 * abs(s1)


# --- changeover ---------------------------------------------------------------
def _K(r, rc, m=1e-5, eps2=0.0):
    D = r * r * r  # eps2 = 0
    return O.kick_weight(r * r, eps2, m, rc) * D / m, O.close_weight(r * r, eps2, m, rc) * D / m


def test_changeover_identities():
    rc = 2.0
    kin, kout = _K(0.1 * rc, rc)
    assert kin == 0.0 and kout == 1.0                       # K(0) = 0
    k1, c1 = _K(rc, rc)
    assert k1 == 1.0 and c1 == 0.0                          # K(1) = 1
    assert _K(0.05 * rc, rc) == (0.0, 1.0)                   # inside 0.1 r_c
    assert _K(3 * rc, rc) == (1.0, 0.0)                      # outside r_c
    for y in np.linspace(0.01, 0.99, 37):
        r = rc * (0.1 + 0.9 * y)
        ry = rc * (0.1 + 0.9 * (1 - y))
        k, c = _K(r, rc)
        k2, _ = _K(ry, rc)
        assert abs(k + c - 1) < 1e-15                       # K + (1-K) = 1
        assert abs(k + k2 - 1) < 1e-14                      # K(y) + K(1-y) = 1
        assert abs(k - y * y / (2 * y * y - 2 * y + 1)) < 1e-14


def test_kick_weight_matches_softened_newton():
    m, eps2 = 3e-5, 1e-8
    r = 50.0
    w = O.kick_weight(r * r, eps2, m, 1.0)                  # far outside r_c: K = 1
    assert w == pytest.approx(m / ((r * r + eps2) * r), rel=1e-15)


def test_hill_radius_of_circular_orbit():
    m = np.array([9.55e-4, 5.15e-5])
    a = np.array([5.2, 30.1])
    x = np.stack([a, 0 * a, 0 * a], 1)
    v = np.stack([0 * a, np.sqrt(1 / a), 0 * a], 1)
    rc = O.hill_radii(1.0, 3.0, m, x, v)
    assert np.allclose(rc, 3 * a * np.cbrt(m / 3), rtol=1e-13)


# --- Bulirsch-Stoer -------------------------------------------------------------
def test_bs_on_pure_kepler_matches_analytic_drift():
    x, v = _state_from_elements(40.0, 0.5, 0.2, 1.0, 2.0, 0.1)
    px = np.array([[1e6, 0, 0]])  # a negligible perturber far outside its changeover zone
    pv = np.array([[0, 1e-3, 0]])
    y, st, nrhs = O.bs_tp(np.concatenate([x, v]), [1e-20], [1.0], px, pv, 0.5)
    xa, va, _, _ = O.kepler_drift(x, v, 1.0, 0.5, 1)
    assert st == 0 and nrhs > 0
    assert np.max(np.abs(y[:3] - xa)) <= 1e-11 * np.linalg.norm(xa)
    assert np.max(np.abs(y[3:] - va)) <= 1e-11 * np.linalg.norm(va)


def test_bs_close_encounter_converges_with_tolerance():
    # TP 0.3 AU from a Neptune-mass body (r_c 2.3 AU): strongly perturbed drift
    x = np.array([30.0, 0.3, 0.0])
    v = np.array([0.0, 0.18, 0.01])
    px = np.array([[30.0, 0.0, 0.0]])
    pv = np.array([[0.0, 0.1826, 0.0]])
    y1, st1, n1 = O.bs_tp(np.concatenate([x, v]), [5.15e-5], [2.33], px, pv, 0.5, tol=1e-10)
    y2, st2, n2 = O.bs_tp(np.concatenate([x, v]), [5.15e-5], [2.33], px, pv, 0.5, tol=1e-13)
    assert st1 == 0 and st2 == 0 and n2 > n1
    assert np.max(np.abs(y1[:3] - y2[:3])) <= 1e-9 * np.linalg.norm(y2[:3])
    # the perturber matters: result differs from the pure Kepler drift
    xa, _, _, _ = O.kepler_drift(x, v, 1.0, 0.5, 1)
    assert np.max(np.abs(y2[:3] - xa)) > 1e-6


def test_bs_spent_halving_budget_is_flagged_and_bounded():
    # A tolerance no error meets (negative; the library refuses it, the
    # oracle's single-drift entry does not): the 24 halvings run out, the
    # drift is flagged (status bit 4) and the rest of the interval is taken in
    # one step.  At most 24 failed attempts, the accepted one and the rest
    # (plus a last sliver from rounding of t), each at most 1 + 2 + ... + 256
    # = 511 right-hand sides; carrying on at h / 2^24 would take millions.
    x = np.array([30.0, 0.3, 0.0])
    v = np.array([0.0, 0.18, 0.01])
    px = np.array([[30.0, 0.0, 0.0]])
    pv = np.array([[0.0, 0.1826, 0.0]])
    y, st, n = O.bs_tp(np.concatenate([x, v]), [5.15e-5], [2.33], px, pv, 0.5, tol=-1.0)
    assert st & 4
    assert n <= 28 * 511, n
    assert np.all(np.isfinite(y))
    # the last row of a full-interval attempt is still a fair estimate
    yc, stc, _ = O.bs_tp(np.concatenate([x, v]), [5.15e-5], [2.33], px, pv, 0.5, tol=1e-13)
    assert stc == 0
    assert np.max(np.abs(y[:3] - yc[:3])) <= 1e-6 * np.linalg.norm(yc[:3])


# --- whole step -------------------------------------------------------------------
class _St:
    def __init__(self, m, mx, mv, tx=None, tv=None):
        self.m = np.asarray(m, float)[None]
        self.mx = np.asarray(mx, float)[None]
        self.mv = np.asarray(mv, float)[None]
        self.tx = np.zeros((0, 3)) if tx is None else np.asarray(tx, float)
        self.tv = np.zeros((0, 3)) if tv is None else np.asarray(tv, float)
        self.tp_offset = np.array([0, len(self.tx)], np.int64)
        self.rc = None


def _earth_mars():
    x1, v1 = _state_from_elements(1.0, 0.017, 0.0, 1.0, 2.0, 0.5)
    x2, v2 = _state_from_elements(1.524, 0.093, 0.032, 4.0, 1.0, 2.5)
    return _St([3e-6, 3.2e-7], [x1, x2], [v1, v2])


def test_energy_error_is_second_order_in_dt():
    st = _earth_mars()
    E0 = O.energy(st.m, st.mx, st.mv)[0]
    errs = []
    for dt, n in ((0.02, 250), (0.01, 500), (0.005, 1000)):
        worst = 0.0
        cur = st
        for _ in range(5):
            out, status, _ = O.run(cur, n // 5, dt)
            cur = _St(st.m[0], out["mx"][0], out["mv"][0])
            cur.rc = out["rc"]
            worst = max(worst, abs(O.energy(st.m, out["mx"], out["mv"])[0] - E0))
        errs.append(worst / abs(E0))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.0 < r1 < 5.0 and 3.0 < r2 < 5.0, errs


def test_earth_mars_energy_bounded_over_1e4_steps():
    st = _earth_mars()
    E0 = O.energy(st.m, st.mx, st.mv)[0]
    out, status, _ = O.run(st, 10000, 1e-3)
    assert status[0] == 0
    assert abs(O.energy(st.m, out["mx"], out["mv"])[0] - E0) <= 1e-10 * abs(E0)


def test_zero_steps_round_trips_democratic_heliocentric_velocities():
    st = _earth_mars()
    tx = np.array([[40.0, 1.0, 0.5]])
    tv = np.array([[0.0, 0.15, 0.01]])
    s2 = _St(st.m[0], st.mx[0], st.mv[0], tx, tv)
    out, _, _ = O.run(s2, 0, 0.01)
    assert np.array_equal(out["mx"], s2.mx)
    assert np.max(np.abs(out["mv"] - s2.mv)) <= 2e-16 * np.max(np.abs(s2.mv))
    assert np.max(np.abs(out["tv"] - tv)) <= 2e-16 * np.max(np.abs(tv))


def test_encounter_check_threshold():
    """A TP 1e-9 inside / outside the changeover radius of a massive body is /
    is not drifted by BS (the jump shifts both bodies by the same vector)."""
    m = 5.15e-5
    xb = np.array([30.0, 0.0, 0.0])
    vb = np.array([0.0, math.sqrt(1 / 30.0), 0.0])
    rc = 2.0
    for f, expect in ((1 - 1e-9, 1), (1 + 1e-9, 0)):
        st = _St([m], [xb], [vb], [xb + np.array([0.0, 0.0, rc * f])], [vb])
        st.rc = np.array([[rc]])
        _, _, stats = O.run(st, 1, 1e-3)
        assert stats["bs_steps"] == expect


def test_tp_is_massless():
    """Adding TPs does not change the massive bodies (bit for bit)."""
    st = _earth_mars()
    a, _, _ = O.run(st, 50, 0.01)
    s2 = _St(st.m[0], st.mx[0], st.mv[0], [[3.0, 0.2, 0.0], [1.2, 0.1, 0.0]], [[0.0, 0.57, 0.0], [0.0, 0.9, 0.01]])
    s2.rc = a["rc"]
    st.rc = a["rc"]
    a, _, _ = O.run(st, 50, 0.01)
    b, _, _ = O.run(s2, 50, 0.01)
    assert np.array_equal(a["mx"], b["mx"]) and np.array_equal(a["mv"], b["mv"])
