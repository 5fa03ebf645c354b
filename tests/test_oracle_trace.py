"""Pins of the oracle's trace definitions, the checker of the library's trace
logger: the Kepler-trace sample (features and the target c = (s*/s_series -
1)/tau^3) and the encounter-trace hit masks (orc_trace_hits)."""
import math

import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import encounter_system
from test_oracle import _state_from_elements


def _c_limit(x, v, mu=1.0):
    """The tau -> 0 limit of the Kepler-trace target, from the fourth-order
    reversion of t(s) = r0 s + eta0 s^2/2 + zeta0 s^3/6 - eta0 beta s^4/24 + ...
    (the universal-variable time equation with c2, c3 expanded): writing
    u = t/r0 = s + p s^2 + q s^3 + w s^4, s = u - p u^2 + (2p^2 - q) u^3 +
    (-5p^3 + 5pq - w) u^4 + ..., and the third-order series leaves the last
    term, so c -> (-5p^3 + 5pq - w) (r0^3/mu)^(1/2)."""
    r0 = float(np.linalg.norm(x))
    eta0 = float(np.dot(x, v))
    beta = 2 * mu / r0 - float(np.dot(v, v))
    zeta0 = mu - beta * r0
    p, q, w = eta0 / (2 * r0), zeta0 / (6 * r0), -eta0 * beta / (24 * r0)
    return (-5 * p ** 3 + 5 * p * q - w) * math.sqrt(r0 ** 3 / mu)


@pytest.mark.parametrize("a,e,nu", [(150.0, 0.7, 1.0), (30.0, 0.1, 2.0), (5.2, 0.048, 0.5), (600.0, 0.93, 0.3)])
def test_kepler_sample_target_tends_to_the_series_remainder(a, e, nu):
    x, v = _state_from_elements(a, e, 0.3, 0.2, 0.1, nu)
    r0 = float(np.linalg.norm(x))
    c0 = _c_limit(x, v)
    errs = []
    for tau in (0.16, 0.08, 0.04, 0.02):
        dt = tau * math.sqrt(r0 ** 3)
        s = O.kepler_sample(x, v, 1.0, dt)
        assert s is not None
        errs.append(abs(float(s[1]) - c0))
    # c(tau) = c0 + O(tau): the error halves with tau (a wrong series
    # coefficient or a wrong root leaves c off by O(1) or growing as tau -> 0)
    for e1, e2 in zip(errs, errs[1:]):
        assert e2 < 0.6 * e1
    assert errs[-1] < 0.05 * abs(c0) + 1e-6


def test_kepler_sample_of_a_circular_orbit_is_exact_series():
    # r = r0 throughout, so s = t / r0 exactly: the series guess is the root
    for r in (5.0, 30.0, 300.0):
        x, v = np.array([r, 0.0, 0.0]), np.array([0.0, math.sqrt(1.0 / r), 0.0])
        f, c = O.kepler_sample(x, v, 1.0, 0.05 * math.sqrt(r ** 3))
        assert abs(float(c)) < 1e-6
        assert abs(float(f[1])) < 1e-6 and float(f[2]) == 0.0 and abs(float(f[3]) - 1.0) < 1e-6


@pytest.mark.parametrize("a,e", [(150.0, 0.7), (5.2, 0.05), (600.0, 0.93)])
def test_kepler_sample_features_at_pericentre(a, e):
    # at pericentre r0 = a (1 - e), eta0 = 0, beta = mu / a: zeta0 / mu = e,
    # eta0 / sqrt(mu r0) = 0, r0 beta / mu = 1 - e; doubling dt adds 1/4 to f0
    x, v = _state_from_elements(a, e, 0.4, 1.0, 2.0, 0.0)
    f1, _ = O.kepler_sample(x, v, 1.0, 0.5)
    f2, _ = O.kepler_sample(x, v, 1.0, 1.0)
    assert abs(float(f1[1]) - e) < 1e-5 and abs(float(f1[2])) < 1e-5 and abs(float(f1[3]) - (1 - e)) < 1e-5
    assert abs((float(f2[0]) - float(f1[0])) - 0.25) < 1e-6
    r0 = a * (1 - e)
    assert abs(float(f1[0]) - (math.log2(0.5 / math.sqrt(r0 ** 3)) + 6) / 4) < 1e-5


def test_kepler_sample_takes_no_unbound_orbit():
    x, v = _state_from_elements(-20.0, 1.5, 0.4, 0.1, 0.2, 0.5)
    assert O.kepler_sample(x, v, 1.0, 0.5) is None
    x = np.array([10.0, 0.0, 0.0])
    v_esc = math.sqrt(2.0 / 10.0)
    assert O.kepler_sample(x, np.array([0.0, v_esc * (1 + 1e-9), 0.0]), 1.0, 0.5) is None  # just unbound
    assert O.kepler_sample(x, np.array([0.0, v_esc * (1 - 1e-9), 0.0]), 1.0, 0.5) is not None  # just bound


def test_trace_without_encounters_is_the_plain_run():
    # no exact check fires: every TP takes its Kepler drift in both, bit for bit
    st = W.p9_single(64)
    hits = O.trace_hits(st, 20, 0.5)
    assert not hits.any()
    ref, _, stats = O.run(st, 20, 0.5)
    assert stats["bs_steps"] == 0
    S, NM = st.m.shape
    m, rc = np.ascontiguousarray(st.m), O.hill_radii(1.0, 3.0, st.m, st.mx, st.mv).reshape(S, NM)
    # trace_hits advances a copy: run it again through the raw call to read the state
    import ctypes as C
    mx, mv = st.mx.astype(float).copy(order="C"), st.mv.astype(float).copy(order="C")
    t = [np.ascontiguousarray(st.tx[:, d], float).copy() for d in range(3)] + \
        [np.ascontiguousarray(st.tv[:, d], float).copy() for d in range(3)]
    h = np.zeros(st.n_tp, np.uint32)
    p = O.OrcParams(0.5, 1.0, 1e-8, 1e-12, 1, 1)
    off = np.ascontiguousarray(st.tp_offset, np.int64)
    O.lib().orc_trace_hits(C.byref(p), S, NM, off.ctypes.data_as(C.POINTER(C.c_int64)), O._d(m), O._d(rc),
                           O._d(mx), O._d(mv), *[O._d(a) for a in t], 20, h.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert np.array_equal(np.stack(t[:3], 1), ref["tx"]) and np.array_equal(np.stack(t[3:], 1), ref["tv"])
    assert np.array_equal(mx, ref["mx"]) and np.array_equal(mv, ref["mv"])


def test_trace_hits_mark_the_bodies_within_their_changeover_radius():
    st = encounter_system(n_tp=400, seed=9)
    S, NM = st.m.shape
    rc = O.hill_radii(1.0, 3.0, st.m, st.mx, st.mv).reshape(S, NM)[0]
    h1 = O.trace_hits(st, 1, 0.5)
    # one step: the check sees the step-start separations shifted by one
    # common jump, so TPs clearly inside or outside (10 %) are determined
    d = np.linalg.norm(st.tx[:, None, :] - st.mx[0][None, :, :], axis=2) / rc[None, :]
    bits = (h1[:, None] >> np.arange(NM, dtype=np.uint32)[None, :]) & 1
    inside, outside = d < 0.9, d > 1.1
    assert inside.sum() > 10 and outside.sum() > 100
    assert np.all(bits[inside] == 1) and np.all(bits[outside] == 0)
    # more steps only add bits
    h8 = O.trace_hits(st, 8, 0.5)
    assert np.all((h1 & ~h8) == 0) and int(np.count_nonzero(h8)) >= int(np.count_nonzero(h1))
