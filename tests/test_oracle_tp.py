"""Pins of the oracle's test-particle side (no GPU): the kick acceleration of a
test particle, the close-encounter drift with perturbers on their Kepler
orbits, the order of the phases of one step, and the long-run behaviour in the
circular restricted three-body problem.  Each is checked against something the
oracle does not share: direct sums in extended precision written from the
changeover formula of src/sim.cuh:166-186,327-382, an adaptive Runge-Kutta
solve (scipy's DOP853) of the same equations of motion, exact symmetries, and
the Jacobi constant of the restricted problem.  A dropped term, a wrong sign
or index, a transposed operand or a phase taken in the wrong order fails one
of them."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import mint
import oracle as O


def _K(y):
    """Mercury's changeover polynomial, clamped (src/sim.cuh:166-186)."""
    y = min(max(y, 0.0), 1.0)
    return y * y / (2 * y * y - 2 * y + 1)


def _K_ld(y):
    y = min(max(y, np.longdouble(0)), np.longdouble(1))
    return y * y / (2 * y * y - 2 * y + 1)


def _accel_ld(m, rc, X, x, eps2):
    """sum_j K_j m_j d_j / ((r_j^2 + eps^2) r_j), d_j = X_j - x, in long double,
    each term formed separately and the terms summed at the end."""
    X = np.asarray(X, np.longdouble)
    x = np.asarray(x, np.longdouble)
    terms = []
    for j in range(len(m)):
        d = X[j] - x
        r = np.sqrt(np.sum(d * d))
        y = (r - np.longdouble(0.1) * np.longdouble(rc[j])) / (np.longdouble(0.9) * np.longdouble(rc[j]))
        terms.append(_K_ld(y) * np.longdouble(m[j]) * d / ((r * r + np.longdouble(eps2)) * r))
    terms = np.array(terms)
    return terms.sum(axis=0), np.abs(terms).sum(axis=0)


def _placed_bodies(nm, seed, regimes):
    """A TP at x and nm bodies at chosen distances r_j = f_j rc_j in random
    directions: f < 0.1 puts the pair inside the inner changeover radius
    (K = 0), 0.1 < f < 1 in the changeover band, f > 1 outside (K = 1)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=3) * 3.0 + np.array([20.0, -5.0, 1.0])
    rc = rng.uniform(0.5, 3.0, nm)
    m = rng.uniform(1e-6, 1e-3, nm)
    f = np.array([regimes[j % len(regimes)] for j in range(nm)]) * rng.uniform(0.9, 1.1, nm)
    u = rng.normal(size=(nm, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    X = x + (f * rc)[:, None] * u
    return m, rc, X, x


@pytest.mark.parametrize("nm", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("regimes", [(0.05,), (0.5,), (3.0,), (0.05, 0.5, 3.0), (0.3, 0.8, 0.11, 20.0)])
def test_tp_accel_matches_long_double_direct_sum(nm, regimes):
    m, rc, X, x = _placed_bodies(nm, 100 * nm + len(regimes), regimes)
    eps2 = 1e-8
    a = O.accel_tp(m, rc, X, x, eps2)
    ref, mag = _accel_ld(m, rc, X, x, eps2)
    tol = 1e-14 * mag.astype(np.float64) + 1e-300
    assert np.all(np.abs(a - ref.astype(np.float64)) <= tol), (a, ref)


def test_tp_accel_regimes_are_the_changeover_limits():
    # every pair inside 0.1 r_c: no kick at all; every pair outside r_c: the
    # full softened Newtonian pull of each body, body by body
    m, rc, X, x = _placed_bodies(5, 7, (0.05,))
    assert np.array_equal(O.accel_tp(m, rc, X, x), np.zeros(3))
    m, rc, X, x = _placed_bodies(5, 8, (1.5, 4.0, 30.0))
    eps2 = 1e-8
    a = O.accel_tp(m, rc, X, x, eps2)
    d = X - x
    r = np.linalg.norm(d, axis=1)
    newton = (m / ((r * r + eps2) * r))[:, None] * d
    assert np.allclose(a, newton.sum(axis=0), rtol=1e-14, atol=0)
    # one body in the band: the pull is K(y) times the Newtonian one, along d
    m, rc, X, x = _placed_bodies(1, 9, (0.55,))
    d = X[0] - x
    r = float(np.linalg.norm(d))
    y = (r - 0.1 * rc[0]) / (0.9 * rc[0])
    a = O.accel_tp(m, rc, X, x, 0.0)
    assert np.allclose(a, _K(y) * m[0] * d / r ** 3, rtol=1e-14, atol=0)
    assert 0.0 < _K(y) < 1.0


def test_tp_accel_scales_exactly_with_the_masses_and_rotates_with_the_bodies():
    m, rc, X, x = _placed_bodies(6, 11, (0.05, 0.4, 0.9, 2.0))
    a = O.accel_tp(m, rc, X, x)
    # doubling every mass doubles every weight's numerator, nothing else
    assert np.array_equal(O.accel_tp(2.0 * m, rc, X, x), 2.0 * a)
    # a rotation of the whole configuration rotates the acceleration (a
    # transposed or mixed component would not)
    rng = np.random.default_rng(3)
    Qm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    ar = O.accel_tp(m, rc, X @ Qm.T, Qm @ x)
    assert np.allclose(ar, Qm @ a, rtol=0, atol=1e-13 * np.linalg.norm(a))
    # the bodies' order only changes the rounding
    p = np.array([3, 0, 5, 1, 4, 2])
    assert np.allclose(O.accel_tp(m[p], rc[p], X[p], x), a, rtol=0, atol=1e-14 * np.linalg.norm(a))


# --- the close-encounter drift with moving perturbers --------------------------
def _tp_pert_rhs(m, rc, eps2, mu):
    """TP: unsoftened Sun plus (1 - K)-weighted softened pulls of the
    perturbers; perturbers: Kepler orbits about the Sun (massless among
    themselves during the drift).  State: [TP (x, v), perturber j (x, v) ...]."""
    nm = len(m)

    def f(t, s):
        out = np.empty_like(s)
        for b in range(nm + 1):
            p = s[6 * b:6 * b + 3]
            out[6 * b:6 * b + 3] = s[6 * b + 3:6 * b + 6]
            out[6 * b + 3:6 * b + 6] = -mu * p / np.linalg.norm(p) ** 3
        x = s[0:3]
        for j in range(nm):
            d = s[6 * (j + 1):6 * (j + 1) + 3] - x
            r = np.linalg.norm(d)
            y = (r - 0.1 * rc[j]) / (0.9 * rc[j])
            out[3:6] += (1.0 - _K(y)) * m[j] * d / ((r * r + eps2) * r)
        return out

    return f


def _encounter_config(seed, frac):
    """Sun (mu = 1), two planets, and a TP at frac r_c of the first planet,
    moving slowly relative to it."""
    rng = np.random.default_rng(seed)
    m = np.array([1e-3, 3e-4])
    rc = np.array([0.21, 0.22])
    ph = rng.uniform(0, 2 * np.pi, 2)
    a = np.array([1.0, 1.6])
    px = np.stack([a * np.cos(ph), a * np.sin(ph), [0.0, 0.02]], 1)
    vc = np.sqrt(1.0 / a)
    pv = np.stack([-vc * np.sin(ph), vc * np.cos(ph), [0.0, 0.0]], 1)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    tx = px[0] + frac * rc[0] * u
    tv = pv[0] + rng.normal(size=3) * 0.05
    return m, rc, px, pv, tx, tv


@pytest.mark.parametrize("frac", [0.04, 0.35, 0.8])
def test_tp_bs_drift_matches_an_ode_solve_with_moving_perturbers(frac):
    m, rc, px, pv, tx, tv = _encounter_config(5, frac)
    h, eps2 = 0.01, 1e-8
    y, st, nrhs = O.bs_tp(np.concatenate([tx, tv]), m, rc, px.ravel(), pv.ravel(), h, mu=1.0, eps2=eps2, tol=1e-13)
    assert st == 0 and nrhs > 0
    s0 = np.concatenate([tx, tv] + [np.concatenate([px[j], pv[j]]) for j in range(2)])
    sol = solve_ivp(_tp_pert_rhs(m, rc, eps2, 1.0), (0, h), s0, method="DOP853", rtol=1e-13, atol=1e-16)
    ref = sol.y[:6, -1]
    assert np.max(np.abs(y[:3] - ref[:3])) < 2e-12
    assert np.max(np.abs(y[3:] - ref[3:])) < 1e-10
    # frozen perturbers would give a different answer at this step length
    frozen = solve_ivp(lambda t, s: _tp_pert_rhs(m, rc, eps2, 1.0)(t, np.concatenate([s, s0[6:]]))[:6],
                       (0, h), s0[:6], method="DOP853", rtol=1e-13, atol=1e-16).y[:, -1]
    assert np.max(np.abs(frozen[3:] - ref[3:])) > 100 * np.max(np.abs(y[3:] - ref[3:]))


def _pair_kick(m, rc, X, i, eps2):
    a = np.zeros(3)
    for k in range(len(m)):
        if k == i:
            continue
        d = X[k] - X[i]
        r = np.linalg.norm(d)
        rcp = max(rc[i], rc[k])
        a += _K((r - 0.1 * rcp) / (0.9 * rcp)) * m[k] * d / ((r * r + eps2) * r)
    return a


def _tp_kick(m, rc, X, x, eps2):
    a = np.zeros(3)
    for j in range(len(m)):
        d = X[j] - x
        r = np.linalg.norm(d)
        a += _K((r - 0.1 * rc[j]) / (0.9 * rc[j])) * m[j] * d / ((r * r + eps2) * r)
    return a


def _one_step_split(m, rc, px, pv, tx, tv, h, eps2, M=1.0, order="reference"):
    """One Mercurius step written from its definition (src/sim.cuh:679-742 with
    the changeover of :166-186): heliocentric to democratic-heliocentric
    velocities, kick h/2, jump h/2 with the momentum after the kick, drift h
    (an ODE solve: the TP in its encounter, the planets on Kepler orbits), jump
    h/2 with the momentum after the drift, kick h/2 with the new positions, and
    back to heliocentric velocities.  `order` plants a wrong phase order in
    this independent restatement (never in the oracle): "late_jump" takes the
    first jump after the drift, "stale_q2" jumps the second time with the
    momentum from before the drift, "early_kick" takes the closing kick before
    the drift.  (A jump moves every body by the same vector, so it commutes
    with the kicks, which see separations only; these orders do not.)"""
    hh = 0.5 * h
    nm = len(m)
    sh = (m[:, None] * pv).sum(0) / (M + m.sum())
    W = pv - sh
    w = tv - sh
    X, x = px, tx
    W = W + hh * np.array([_pair_kick(m, rc, X, i, eps2) for i in range(nm)])
    w = w + hh * _tp_kick(m, rc, X, x, eps2)
    if order == "early_kick":
        W = W + hh * np.array([_pair_kick(m, rc, X, i, eps2) for i in range(nm)])
        w = w + hh * _tp_kick(m, rc, X, x, eps2)
    Q1 = (m[:, None] * W).sum(0) / M
    if order != "late_jump":
        X, x = X + hh * Q1, x + hh * Q1
    assert any(np.sum((X[j] - x) ** 2) < rc[j] ** 2 for j in range(nm)), "the TP must be flagged"
    s0 = np.concatenate([x, w] + [np.concatenate([X[j], W[j]]) for j in range(nm)])
    sol = solve_ivp(_tp_pert_rhs(m, rc, eps2, M), (0, h), s0, method="DOP853", rtol=1e-13, atol=1e-16)
    s1 = sol.y[:, -1]
    x, w = s1[0:3], s1[3:6]
    X = np.array([s1[6 * (j + 1):6 * (j + 1) + 3] for j in range(nm)])
    W = np.array([s1[6 * (j + 1) + 3:6 * (j + 2)] for j in range(nm)])
    if order == "late_jump":
        X, x = X + hh * Q1, x + hh * Q1
    Q2 = Q1 if order == "stale_q2" else (m[:, None] * W).sum(0) / M
    X, x = X + hh * Q2, x + hh * Q2
    if order != "early_kick":
        W = W + hh * np.array([_pair_kick(m, rc, X, i, eps2) for i in range(nm)])
        w = w + hh * _tp_kick(m, rc, X, x, eps2)
    back = (m[:, None] * W).sum(0) / M
    return x, w + back, X, W + back


def _state(m, rc, px, pv, tx, tv):
    return mint.State(np.array([0, 1], np.int64), m[None].copy(), px[None].copy(), pv[None].copy(),
                      tx[None].copy(), tv[None].copy(), rc=rc[None].copy())


@pytest.mark.parametrize("frac,seed", [(0.04, 1), (0.35, 2), (0.8, 3), (0.6, 4)])
def test_tp_step_in_a_changeover_zone_matches_the_split_solved_independently(frac, seed):
    m, rc, px, pv, tx, tv = _encounter_config(seed, frac)
    h, eps2 = 0.005, 1e-8
    out, status, stats = O.run(_state(m, rc, px, pv, tx, tv), 1, h, softening2=eps2, bs_tol=1e-13)
    assert stats["bs_steps"] == 1 and stats["kepler_solves"] == 0 and status[0] == 0
    x1, v1, X1, V1 = _one_step_split(m, rc, px, pv, tx, tv, h, eps2)
    assert np.max(np.abs(out["tx"][0] - x1)) < 2e-12
    assert np.max(np.abs(out["tv"][0] - v1)) < 2e-10
    assert np.max(np.abs(out["mx"][0] - X1)) < 1e-13
    assert np.max(np.abs(out["mv"][0] - V1)) < 1e-12


@pytest.mark.parametrize("order", ["late_jump", "stale_q2", "early_kick"])
def test_tp_step_phase_order_is_resolved_by_the_independent_split(order):
    # a wrong phase order planted in the independent split lands far outside
    # the agreement the oracle reaches with the right one, so the comparison
    # above pins the order of the phases, not only their formulas
    m, rc, px, pv, tx, tv = _encounter_config(2, 0.35)
    h, eps2 = 0.005, 1e-8
    out, _, _ = O.run(_state(m, rc, px, pv, tx, tv), 1, h, softening2=eps2, bs_tol=1e-13)
    x1, v1, _, _ = _one_step_split(m, rc, px, pv, tx, tv, h, eps2)
    xw, vw, _, _ = _one_step_split(m, rc, px, pv, tx, tv, h, eps2, order=order)
    ok = max(np.max(np.abs(out["tx"][0] - x1)), np.max(np.abs(out["tv"][0] - v1)))
    bad = max(np.max(np.abs(out["tx"][0] - xw)), np.max(np.abs(out["tv"][0] - vw)))
    assert bad > 1e3 * ok, (ok, bad)


# --- the circular restricted three-body problem --------------------------------
def _cr3bp(dt, nsteps, a_tp, e_tp=0.0, seed=0, m_p=1e-3):
    """Sun (M = 1) and one planet on a circular orbit (a = 1) with a test
    particle; returns the Jacobi constant after every `nsteps // 50` steps and
    the oracle's run statistics."""
    rng = np.random.default_rng(seed)
    M = 1.0
    n = math.sqrt((M + m_p))
    px = np.array([[1.0, 0.0, 0.0]])
    pv = np.array([[0.0, n, 0.0]])
    ph = rng.uniform(0.5, 2 * np.pi - 0.5)
    r = a_tp * (1 - e_tp)
    vq = math.sqrt(M * (1 + e_tp) / r)
    tx = np.array([[r * math.cos(ph), r * math.sin(ph), 0.0]])
    tv = np.array([[-vq * math.sin(ph), vq * math.cos(ph), 0.0]])
    rc = np.array([3.0 * (m_p / 3.0) ** (1 / 3)])
    st = mint.State(np.array([0, 1], np.int64), np.array([[m_p]]), px[None].copy(), pv[None].copy(), tx, tv,
                    rc=rc[None])

    def jacobi(s):
        xp, vp = s.mx[0, 0], s.mv[0, 0]
        x, v = s.tx[0], s.tv[0]
        xb = x - m_p * xp / (M + m_p)
        vb = v - m_p * vp / (M + m_p)
        r1 = np.linalg.norm(x)
        r2 = np.linalg.norm(x - xp)
        return 2 * (M / r1 + m_p / r2) - vb @ vb + 2 * n * (xb[0] * vb[1] - xb[1] * vb[0])

    C = [jacobi(st)]
    chunk = nsteps // 50
    tot = {"bs_steps": 0, "kepler_solves": 0}
    for _ in range(50):
        out, status, stats = O.run(st, chunk, dt, softening2=0.0, bs_tol=1e-12)
        assert status[0] == 0
        for k in tot:
            tot[k] += stats[k]
        st = mint.State(st.tp_offset, st.m, out["mx"], out["mv"], out["tx"], out["tv"], rc=st.rc)
        C.append(jacobi(st))
    return np.array(C), tot


def test_jacobi_constant_is_bounded_and_second_order_without_encounters():
    errs = []
    for dt in (0.02, 0.01):
        nsteps = int(round(200.0 / dt))  # 32 orbits of the planet
        C, tot = _cr3bp(dt, nsteps, a_tp=1.55, seed=4)
        assert tot["bs_steps"] == 0
        errs.append(np.max(np.abs(C - C[0])) / abs(C[0]))
    assert errs[0] < 1e-6
    ratio = errs[0] / errs[1]
    assert 3.2 < ratio < 5.0, (errs, ratio)


def test_jacobi_constant_stays_second_order_through_close_encounters():
    # an orbit crossing the planet's: the changeover hands the encounters to
    # BS, and the error of the Jacobi constant still falls fourfold per halving
    # of the step (a changeover weight applied twice, or missing from either
    # side of the split, would leave a first-order or constant error)
    errs = []
    for dt in (0.01, 0.005):
        C, tot = _cr3bp(dt, int(round(200.0 / dt)), a_tp=1.25, e_tp=0.25, seed=1)
        assert tot["bs_steps"] > 0.02 * tot["kepler_solves"]
        errs.append(np.max(np.abs(C - C[0])) / abs(C[0]))
    assert errs[0] < 1e-6
    assert 3.2 < errs[0] / errs[1] < 5.0, errs
