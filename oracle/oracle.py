"""ctypes view of the CPU oracle (oracle/build/libmint_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the library or its host mirror.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libmint_oracle.so")
_DP = C.POINTER(C.c_double)


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("central_mass", C.c_double), ("softening2", C.c_double),
                ("bs_tol", C.c_double), ("guess", C.c_int32), ("threads", C.c_int32)]


class OrcStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("tp_steps", "kepler_solves", "newton_iters", "bs_steps", "bs_rhs_evals",
                                         "massive_groups", "massive_group_bodies", "overlaps",
                                         "massive_kepler_solves", "massive_newton_iters")]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built at {LIB_PATH}; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_stumpff.argtypes = [C.c_double, _DP]
        L.orc_kepler_drift.argtypes = [_DP, _DP, C.c_double, C.c_double, C.c_int, _DP]
        L.orc_kepler_drift.restype = C.c_int
        L.orc_kick_weight.argtypes = [C.c_double] * 4
        L.orc_kick_weight.restype = C.c_double
        L.orc_close_weight.argtypes = [C.c_double] * 4
        L.orc_close_weight.restype = C.c_double
        L.orc_hill_radii.argtypes = [C.c_int64, C.c_double, C.c_double, _DP, _DP, _DP, _DP]
        L.orc_bs_tp.argtypes = [_DP, C.c_int, _DP, _DP, _DP, _DP, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.POINTER(C.c_int64)]
        L.orc_bs_tp.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(OrcParams), C.c_int32, C.c_int32, C.POINTER(C.c_int64), _DP, _DP,
                              _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, C.c_int64, C.POINTER(C.c_uint32),
                              C.POINTER(OrcStats)]
        L.orc_run.restype = C.c_int
        L.orc_energy.argtypes = [C.c_int32, C.c_int32, _DP, _DP, _DP, C.c_double, C.c_double, _DP]
        L.orc_accel_massive.argtypes = [C.c_int32, _DP, _DP, _DP, C.c_double, _DP]
        L.orc_encounter_groups.argtypes = [C.c_int32, _DP, _DP, _DP, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.orc_encounter_groups.restype = C.c_int
        L.orc_accel_tp.argtypes = [C.c_int32, _DP, _DP, _DP, _DP, C.c_double, _DP]
        L.orc_accel_tp.restype = C.c_int
        L.orc_trace_hits.argtypes = [C.POINTER(OrcParams), C.c_int32, C.c_int32, C.POINTER(C.c_int64), _DP, _DP,
                                     _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, C.c_int64, C.POINTER(C.c_uint32)]
        L.orc_trace_hits.restype = C.c_int
        L.orc_kepler_sample.argtypes = [_DP, _DP, C.c_double, C.c_double, C.POINTER(C.c_float),
                                        C.POINTER(C.c_float)]
        L.orc_kepler_sample.restype = C.c_int
        _L = L
    return _L


def _d(a):
    return a.ctypes.data_as(_DP)


def stumpff(z):
    c = np.empty(4)
    lib().orc_stumpff(float(z), _d(c))
    return c


def kepler_drift(x, v, mu, dt, guess=1):
    x = np.array(x, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    s = np.zeros(1)
    it = lib().orc_kepler_drift(_d(x), _d(v), mu, dt, guess, _d(s))
    return x, v, it, float(s[0])


def kick_weight(r2, eps2, m, rc):
    return lib().orc_kick_weight(r2, eps2, m, rc)


def close_weight(r2, eps2, m, rc):
    return lib().orc_close_weight(r2, eps2, m, rc)


def hill_radii(M, n_hill, m, x, v):
    m = np.ascontiguousarray(m, np.float64).ravel()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    rc = np.empty(len(m))
    lib().orc_hill_radii(len(m), M, n_hill, _d(m), _d(x), _d(v), _d(rc))
    return rc


def bs_tp(y6, m, rc, px, pv, h, mu=1.0, eps2=1e-8, tol=1e-12):
    y = np.array(y6, dtype=np.float64)
    m = np.ascontiguousarray(m, np.float64)
    rc = np.ascontiguousarray(rc, np.float64)
    px = np.ascontiguousarray(px, np.float64)
    pv = np.ascontiguousarray(pv, np.float64)
    n = C.c_int64()
    st = lib().orc_bs_tp(_d(y), len(m), _d(m), _d(rc), _d(px), _d(pv), h, mu, eps2, tol, C.byref(n))
    return y, st, n.value


def run(state, nsteps, dt, central_mass=1.0, softening2=1e-8, bs_tol=1e-12, guess=1, threads=1, n_hill=3.0,
        radii=None):
    """Integrate a copy of `state` (a mint.State-like object); returns
    (new_state_arrays dict, status, stats).  radii: optional physical radii
    [S, NM] (overlaps are counted, nothing merges)."""
    S, NM = state.m.shape
    off = np.ascontiguousarray(state.tp_offset, np.int64)
    m = np.ascontiguousarray(state.m, np.float64)
    rc = state.rc if state.rc is not None else hill_radii(central_mass, n_hill, state.m, state.mx, state.mv).reshape(S, NM)
    rc = np.ascontiguousarray(rc, np.float64)
    mx = np.array(state.mx, dtype=np.float64, order="C")
    mv = np.array(state.mv, dtype=np.float64, order="C")
    t = [np.ascontiguousarray(state.tx[:, d], np.float64).copy() for d in range(3)] + \
        [np.ascontiguousarray(state.tv[:, d], np.float64).copy() for d in range(3)]
    status = np.zeros(S, np.uint32)
    stats = OrcStats()
    p = OrcParams(dt, central_mass, softening2, bs_tol, guess, threads)
    rad = None if radii is None else np.ascontiguousarray(radii, np.float64).reshape(S, NM)
    rc_ = lib().orc_run(C.byref(p), S, NM, off.ctypes.data_as(C.POINTER(C.c_int64)), _d(m), _d(rc),
                        None if rad is None else _d(rad), _d(mx), _d(mv), *[_d(a) for a in t], int(nsteps),
                        status.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(stats))
    if rc_ != 0:
        raise ValueError("oracle: test particles need at most 8 massive bodies")
    out = dict(mx=mx, mv=mv, tx=np.stack(t[:3], 1), tv=np.stack(t[3:], 1), rc=rc)
    return out, status, {n: getattr(stats, n) for n, _ in OrcStats._fields_}


def trace_hits(state, nsteps, dt, central_mass=1.0, softening2=1e-8, bs_tol=1e-12, n_hill=3.0):
    """Per TP, the bitmask of bodies whose exact check fires in `nsteps`
    steps along the TP's Kepler path (the encounter-trace labels)."""
    S, NM = state.m.shape
    off = np.ascontiguousarray(state.tp_offset, np.int64)
    m = np.ascontiguousarray(state.m, np.float64)
    rc = state.rc if state.rc is not None else hill_radii(central_mass, n_hill, state.m, state.mx, state.mv).reshape(S, NM)
    rc = np.ascontiguousarray(rc, np.float64)
    mx = np.array(state.mx, dtype=np.float64, order="C")
    mv = np.array(state.mv, dtype=np.float64, order="C")
    t = [np.ascontiguousarray(state.tx[:, d], np.float64).copy() for d in range(3)] + \
        [np.ascontiguousarray(state.tv[:, d], np.float64).copy() for d in range(3)]
    hits = np.zeros(int(off[-1]), np.uint32)
    p = OrcParams(dt, central_mass, softening2, bs_tol, 1, 1)
    lib().orc_trace_hits(C.byref(p), S, NM, off.ctypes.data_as(C.POINTER(C.c_int64)), _d(m), _d(rc), _d(mx), _d(mv),
                         *[_d(a) for a in t], int(nsteps), hits.ctypes.data_as(C.POINTER(C.c_uint32)))
    return hits


def kepler_sample(x, v, mu, dt):
    """(features[4], target) of the Kepler trace for a drift from the DH
    state (x, v), or None where the trace takes no sample."""
    x = np.ascontiguousarray(x, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    f = np.zeros(4, np.float32)
    c = C.c_float()
    if not lib().orc_kepler_sample(_d(x), _d(v), mu, dt, f.ctypes.data_as(C.POINTER(C.c_float)), C.byref(c)):
        return None
    return f, np.float32(c.value)


def energy(m, mx, mv, M=1.0, eps2=1e-8):
    m = np.ascontiguousarray(m, np.float64)
    S, NM = m.shape
    E = np.empty(S)
    lib().orc_energy(S, NM, _d(m), _d(np.ascontiguousarray(mx, np.float64)), _d(np.ascontiguousarray(mv, np.float64)),
                     M, eps2, _d(E))
    return E


def accel_massive(m, rc, x, eps2=1e-8):
    m = np.ascontiguousarray(m, np.float64).ravel()
    rc = np.ascontiguousarray(rc, np.float64).ravel()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    a = np.empty_like(x)
    lib().orc_accel_massive(len(m), _d(m), _d(rc), _d(x), eps2, _d(a))
    return a


def accel_tp(m, rc, X, x, eps2=1e-8):
    """Kick acceleration of one test particle at x from the massive bodies at X."""
    m = np.ascontiguousarray(m, np.float64).ravel()
    rc = np.ascontiguousarray(rc, np.float64).ravel()
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 3)
    x = np.ascontiguousarray(x, np.float64).ravel()
    a = np.empty(3)
    if lib().orc_accel_tp(len(m), _d(m), _d(rc), _d(X), _d(x), eps2, _d(a)) != 0:
        raise ValueError("oracle: 1..8 massive bodies")
    return a


def encounter_groups(rc, x, radii=None):
    """(gid[NM] with -1 for bodies in no group, number of groups, overlaps)."""
    rc = np.ascontiguousarray(rc, np.float64).ravel()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    rad = None if radii is None else np.ascontiguousarray(radii, np.float64).ravel()
    gid = np.empty(len(rc), np.int32)
    ov = C.c_int64()
    n = lib().orc_encounter_groups(len(rc), _d(rc), None if rad is None else _d(rad), _d(x),
                                   gid.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(ov))
    return gid, n, ov.value
