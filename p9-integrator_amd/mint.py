"""Python view of libmint.so, the MI355X Mercurius step library (C ABI in
include/mint.h).

This is the host-side mirror used by the tests and the benchmark: it owns no
numerics, only marshals numpy arrays through ctypes.  The library must be
built (``make -C p9-integrator_amd``); importing this module raises if it is
missing, and every call that needs the GPU fails loudly without one.

Reference interface mirrored: the ``mint`` driver's flow in src/mercurius.cu:17-263
(config -> upload -> launch per batch -> download), with the element arrays
replaced by a CSR batch of Cartesian states.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmint.so")

FORM_PLAIN, FORM_FLAT, FORM_BATCHED = 0, 1, 2
GUESS_CRUDE, GUESS_SERIES, GUESS_LEARNED = 0, 1, 2
ST_NAN, ST_NEWTON, ST_BS_DEPTH, ST_MASSIVE_ENC, ST_OVERLAP, ST_LIMIT = 1, 2, 4, 8, 16, 32
MAX_NM_TP = 8         # massive bodies per system when the system has test particles
MAX_NM = 2048         # massive bodies per system without test particles
NET_KEPLER, NET_SCORER = 0, 1


class MintError(RuntimeError):
    pass


class _Config(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("central_mass", C.c_double),
        ("n_hill", C.c_double),
        ("softening2", C.c_double),
        ("bs_tol", C.c_double),
        ("form", C.c_int32),
        ("kepler_guess", C.c_int32),
        ("systems_per_launch", C.c_int32),
        ("block_size", C.c_int32),
        ("tps_per_lane", C.c_int32),
        ("chunk_steps", C.c_int32),
        ("use_scorer", C.c_int32),
        ("use_graph", C.c_int32),
        ("device", C.c_int32),
        ("collect_stats", C.c_int32),
        ("many_body", C.c_int32),
    ]


_DP = C.POINTER(C.c_double)


class _State(C.Structure):
    _fields_ = [
        ("n_systems", C.c_int32),
        ("n_massive", C.c_int32),
        ("tp_offset", C.POINTER(C.c_int64)),
        ("m", _DP),
        ("mx", _DP),
        ("mv", _DP),
        ("tx", _DP),
        ("ty", _DP),
        ("tz", _DP),
        ("tvx", _DP),
        ("tvy", _DP),
        ("tvz", _DP),
        ("rc", _DP),
    ]


class _Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("tp_steps", "kepler_solves", "newton_iters", "bs_steps",
                                         "bs_rhs_evals", "deferred", "prebucketed", "idle_lane_steps",
                                         "enc_tp_steps", "massive_groups", "massive_group_bodies",
                                         "overlaps")]


class _Timing(C.Structure):
    _fields_ = [
        ("ms_ephemeris", C.c_double), ("n_ephemeris", C.c_int64),
        ("ms_tp_main", C.c_double), ("n_tp_main", C.c_int64),
        ("ms_encounter", C.c_double), ("n_encounter", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_mb_open", C.c_double), ("n_mb_open", C.c_int64),
        ("ms_mb_drift", C.c_double), ("n_mb_drift", C.c_int64),
        ("ms_mb_close", C.c_double), ("n_mb_close", C.c_int64),
    ]


# symbol -> (restype, argtypes); this list is what include/mint.h declares
_SIGS = {
    "mint_abi_version": (C.c_int32, []),
    "mint_config_default": (None, [C.POINTER(_Config)]),
    "mint_create": (C.c_int, [C.POINTER(_Config), C.POINTER(C.c_void_p)]),
    "mint_destroy": (None, [C.c_void_p]),
    "mint_last_error": (C.c_char_p, [C.c_void_p]),
    "mint_upload": (C.c_int, [C.c_void_p, C.POINTER(_State)]),
    "mint_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "mint_prepare": (C.c_int, [C.c_void_p, C.c_int64]),
    "mint_flops_per_tp_step": (C.c_double, [C.c_int32, C.c_double]),
    "mint_flops_per_kick_pass": (C.c_double, [C.c_int64]),
    "mint_step_noalloc": (C.c_int, [C.c_void_p, C.c_int64]),
    "mint_download": (C.c_int, [C.c_void_p, C.POINTER(_State)]),
    "mint_energy": (C.c_int, [C.c_void_p, _DP]),
    "mint_status": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "mint_get_stats": (C.c_int, [C.c_void_p, C.POINTER(_Stats)]),
    "mint_get_timing": (C.c_int, [C.c_void_p, C.POINTER(_Timing)]),
    "mint_reset_counters": (C.c_int, [C.c_void_p]),
    "mint_set_mode": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "mint_hill_radii": (C.c_int, [C.c_int64, C.c_double, C.c_double, _DP, _DP, _DP, _DP]),
    "mint_selftest_arith": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, _DP, _DP, _DP]),
    "mint_elements_to_cartesian": (C.c_int, [C.c_int64, C.c_double] + [_DP] * 8),
    "mint_cartesian_to_elements": (C.c_int, [C.c_int64, C.c_double] + [_DP] * 8),
    "mint_autotune": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int32)]),
    "mint_launch_shape": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]),
    "mint_trace_kepler": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_int64)]),
    "mint_trace_encounter": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), C.POINTER(C.c_int64)]),
    "mint_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float),
                             C.c_int32, C.c_int32, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "mint_nn_forward": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float)]),
    "mint_nn_loss_grad": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mint_nn_nparams": (C.c_int, [C.c_int32]),
    "mint_nn_get_weights": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_float)]),
    "mint_nn_set_weights": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_float)]),
    "mint_set_score_threshold": (C.c_int, [C.c_void_p, C.c_float]),
    "mint_scorer_skip": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_int32)]),
    "mint_save_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "mint_load_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]),
    "mint_set_radii": (C.c_int, [C.c_void_p, _DP]),
}


ABI_VERSION = 2


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise MintError(f"libmint.so not built at {path}; run `make -C p9-integrator_amd`")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.mint_abi_version() != ABI_VERSION:
        raise MintError(f"{path} has ABI {lib.mint_abi_version()}, this mirror binds ABI {ABI_VERSION}; rebuild")
    return lib


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def _d(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class State:
    """CSR batch of systems: massive bodies (S, Nm, 3) and SoA TPs, heliocentric."""
    tp_offset: np.ndarray          # int64 [S+1]
    m: np.ndarray                  # [S, Nm]
    mx: np.ndarray                 # [S, Nm, 3]
    mv: np.ndarray                 # [S, Nm, 3]
    tx: np.ndarray                 # [N, 3] positions
    tv: np.ndarray                 # [N, 3] velocities
    rc: np.ndarray | None = None   # [S, Nm]
    meta: dict = field(default_factory=dict)

    @property
    def n_systems(self):
        return self.m.shape[0]

    @property
    def n_massive(self):
        return self.m.shape[1]

    @property
    def n_tp(self):
        return int(self.tp_offset[-1])

    def copy(self) -> "State":
        return State(self.tp_offset.copy(), self.m.copy(), self.mx.copy(), self.mv.copy(), self.tx.copy(),
                     self.tv.copy(), None if self.rc is None else self.rc.copy(), dict(self.meta))

    def subset(self, systems) -> "State":
        systems = np.asarray(systems, dtype=np.int64)
        counts = np.diff(self.tp_offset)[systems]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        idx = np.concatenate([np.arange(self.tp_offset[s], self.tp_offset[s + 1]) for s in systems]) \
            if len(systems) else np.zeros(0, np.int64)
        return State(off, self.m[systems].copy(), self.mx[systems].copy(), self.mv[systems].copy(),
                     self.tx[idx].copy(), self.tv[idx].copy(),
                     None if self.rc is None else self.rc[systems].copy(), dict(self.meta))


class _Marshal:
    """Contiguous SoA buffers handed to the library for one call."""

    def __init__(self, st: State, for_output=False):
        self.off = np.ascontiguousarray(st.tp_offset, dtype=np.int64)
        self.m = np.ascontiguousarray(st.m, dtype=np.float64)
        self.mx = np.ascontiguousarray(st.mx, dtype=np.float64)
        self.mv = np.ascontiguousarray(st.mv, dtype=np.float64)
        n = st.n_tp
        if for_output:
            self.t = [np.empty(n) for _ in range(6)]
        else:
            self.t = [np.ascontiguousarray(st.tx[:, d]) for d in range(3)] + \
                     [np.ascontiguousarray(st.tv[:, d]) for d in range(3)]
        self.rc = None if st.rc is None else np.ascontiguousarray(st.rc, dtype=np.float64)
        self.s = _State(st.n_systems, st.n_massive, self.off.ctypes.data_as(C.POINTER(C.c_int64)),
                        _d(self.m), _d(self.mx), _d(self.mv), *[_d(a) for a in self.t], _d(self.rc))


class Mint:
    """One library context (one HIP stream on one device)."""

    def __init__(self, dt=0.8219, central_mass=1.0, n_hill=3.0, softening2=1e-8, bs_tol=1e-12,
                 form=FORM_BATCHED, kepler_guess=GUESS_SERIES, systems_per_launch=4, block_size=0,
                 tps_per_lane=0, chunk_steps=0, use_scorer=0, use_graph=1, device=0, collect_stats=0,
                 many_body=0):
        self.L = lib()
        cfg = _Config()
        self.L.mint_config_default(C.byref(cfg))
        for k, v in dict(dt=dt, central_mass=central_mass, n_hill=n_hill, softening2=softening2, bs_tol=bs_tol,
                         form=form, kepler_guess=kepler_guess, systems_per_launch=systems_per_launch,
                         block_size=block_size, tps_per_lane=tps_per_lane, chunk_steps=chunk_steps,
                         use_scorer=use_scorer, use_graph=use_graph, device=device,
                         collect_stats=collect_stats, many_body=many_body).items():
            setattr(cfg, k, v)
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.L.mint_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise MintError(f"mint_create failed ({rc}): {self.L.mint_last_error(None).decode()}")
        self.h = h
        self._shape = None

    def close(self):
        if getattr(self, "h", None):
            self.L.mint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            raise MintError(f"{what} failed ({rc}): {self.L.mint_last_error(self.h).decode()}")

    def upload(self, st: State):
        mm = _Marshal(st)
        self._chk(self.L.mint_upload(self.h, C.byref(mm.s)), "mint_upload")
        self._shape = (st.n_systems, st.n_massive, st.tp_offset.copy())

    def step(self, nsteps: int):
        self._chk(self.L.mint_step(self.h, int(nsteps)), "mint_step")

    def step_noalloc(self, nsteps: int):
        """mint_step_noalloc: the flat per-step schedule with the library's
        fixed device scratch (no allocation)."""
        self._chk(self.L.mint_step_noalloc(self.h, int(nsteps)), "mint_step_noalloc")

    def prepare(self, nsteps: int):
        """Build the launch graph of an nsteps call ahead of it (host work only)."""
        self._chk(self.L.mint_prepare(self.h, int(nsteps)), "mint_prepare")

    def download(self) -> State:
        S, NM, off = self._shape
        n = int(off[-1])
        st = State(off.copy(), np.empty((S, NM)), np.empty((S, NM, 3)), np.empty((S, NM, 3)),
                   np.empty((n, 3)), np.empty((n, 3)), np.empty((S, NM)))
        mm = _Marshal(st, for_output=True)
        self._chk(self.L.mint_download(self.h, C.byref(mm.s)), "mint_download")
        st.m[:] = mm.m
        st.mx[:] = mm.mx
        st.mv[:] = mm.mv
        st.rc[:] = mm.rc
        for d in range(3):
            st.tx[:, d] = mm.t[d]
            st.tv[:, d] = mm.t[3 + d]
        return st

    def save_checkpoint(self, path, step_index=0):
        self._chk(self.L.mint_save_checkpoint(self.h, path.encode(), int(step_index)), "mint_save_checkpoint")

    def load_checkpoint(self, path) -> int:
        k = C.c_int64()
        self._chk(self.L.mint_load_checkpoint(self.h, path.encode(), C.byref(k)), "mint_load_checkpoint")
        # the shape comes from the file header: magic, int32 S, NM, version, int64 Ntp, step
        hdr = np.fromfile(path, dtype=np.int32, count=5)[2:5]
        n = int(np.fromfile(path, dtype=np.int64, count=1, offset=20)[0])
        off = np.fromfile(path, dtype=np.int64, count=int(hdr[0]) + 1, offset=20 + 16 + 40)
        self._shape = (int(hdr[0]), int(hdr[1]), off)
        assert int(off[-1]) == n
        return k.value

    def set_radii(self, radius):
        """Physical radii [S, NM] (recorded and checked for overlaps; nothing merges)."""
        r = np.ascontiguousarray(radius, np.float64).ravel()
        S, NM = self._shape[0], self._shape[1]
        if r.size != S * NM:
            raise MintError(f"set_radii: need {S * NM} radii, got {r.size}")
        self._chk(self.L.mint_set_radii(self.h, _d(r)), "mint_set_radii")

    def energy(self) -> np.ndarray:
        E = np.empty(self._shape[0])
        self._chk(self.L.mint_energy(self.h, _d(E)), "mint_energy")
        return E

    def status(self) -> np.ndarray:
        s = np.zeros(self._shape[0], dtype=np.uint32)
        self._chk(self.L.mint_status(self.h, s.ctypes.data_as(C.POINTER(C.c_uint32))), "mint_status")
        return s

    def stats(self) -> dict:
        s = _Stats()
        self._chk(self.L.mint_get_stats(self.h, C.byref(s)), "mint_get_stats")
        return {n: getattr(s, n) for n, _ in _Stats._fields_}

    def timing(self) -> dict:
        t = _Timing()
        self._chk(self.L.mint_get_timing(self.h, C.byref(t)), "mint_get_timing")
        return {n: getattr(t, n) for n, _ in _Timing._fields_}

    def reset_counters(self):
        self._chk(self.L.mint_reset_counters(self.h), "mint_reset_counters")

    def set_mode(self, form, kepler_guess, use_scorer=0):
        self._chk(self.L.mint_set_mode(self.h, form, kepler_guess, use_scorer), "mint_set_mode")

    def autotune(self, cache_path=None):
        best = (C.c_int32 * 3)()
        self._chk(self.L.mint_autotune(self.h, cache_path.encode() if cache_path else None, best), "mint_autotune")
        return tuple(best)

    def launch_shape(self, nsteps):
        """(block, TPs per lane, window) that a batched call of nsteps steps uses."""
        sh = (C.c_int32 * 3)()
        self._chk(self.L.mint_launch_shape(self.h, int(nsteps), sh), "mint_launch_shape")
        return tuple(sh)

    # learned parts -------------------------------------------------------
    def trace_kepler(self, nsteps, cap):
        feat = np.empty((cap, 4), np.float32)
        tgt = np.empty(cap, np.float32)
        n = C.c_int64()
        self._chk(self.L.mint_trace_kepler(self.h, nsteps, cap, _f(feat), _f(tgt), C.byref(n)), "mint_trace_kepler")
        return feat[: n.value].copy(), tgt[: n.value].copy()

    def trace_encounter(self, nsteps, cap):
        feat = np.empty((cap, 6), np.float32)
        lab = np.empty(cap, np.float32)
        n = C.c_int64()
        self._chk(self.L.mint_trace_encounter(self.h, nsteps, cap, _f(feat), _f(lab), C.byref(n)),
                  "mint_trace_encounter")
        return feat[: n.value].copy(), lab[: n.value].copy()

    def train(self, which, feat, target, epochs=5, batch=4096, lr=1e-3, seed=1):
        feat = np.ascontiguousarray(feat, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        losses = np.zeros(epochs, np.float32)
        self._chk(self.L.mint_train(self.h, which, len(target), _f(feat), _f(target), epochs, batch, lr, seed,
                                    _f(losses)), "mint_train")
        return losses

    def nn_forward(self, which, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.empty(len(feat), np.float32)
        self._chk(self.L.mint_nn_forward(self.h, which, len(feat), _f(feat), _f(out)), "mint_nn_forward")
        return out

    def nn_loss_grad(self, which, feat, target):
        feat = np.ascontiguousarray(feat, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        g = np.empty(self.L.mint_nn_nparams(which), np.float32)
        loss = C.c_float()
        self._chk(self.L.mint_nn_loss_grad(self.h, which, len(target), _f(feat), _f(target), C.byref(loss), _f(g)),
                  "mint_nn_loss_grad")
        return loss.value, g

    def nn_weights(self, which):
        w = np.empty(self.L.mint_nn_nparams(which), np.float32)
        self._chk(self.L.mint_nn_get_weights(self.h, which, _f(w)), "mint_nn_get_weights")
        return w

    def set_nn_weights(self, which, w):
        w = np.ascontiguousarray(w, np.float32)
        assert len(w) == self.L.mint_nn_nparams(which)
        self._chk(self.L.mint_nn_set_weights(self.h, which, _f(w)), "mint_nn_set_weights")

    def scorer_skip(self, mode=-1):
        """Set the far-pair skip of the scorer (1 where provably neutral, 0
        never, -1 unchanged); returns (interval bound of the logit over the
        far pairs, the cutoff c in units of 8 r_c, whether pairs are skipped)."""
        b = C.c_float()
        cut = C.c_float()
        a = C.c_int32()
        self._chk(self.L.mint_scorer_skip(self.h, int(mode), C.byref(b), C.byref(cut), C.byref(a)),
                  "mint_scorer_skip")
        return b.value, cut.value, bool(a.value)

    def set_score_threshold(self, t):
        self._chk(self.L.mint_set_score_threshold(self.h, float(t)), "mint_set_score_threshold")


# host helpers (no GPU needed) -------------------------------------------------
def flops_per_tp_step(nbar: float, nm: int = 5) -> float:
    """Algorithmic FLOP per test-particle step (include/mint_flops.h)."""
    return float(lib().mint_flops_per_tp_step(int(nm), float(nbar)))


def flops_per_kick_pass(n: int) -> float:
    """Algorithmic FLOP of one massive kick-acceleration pass over n bodies."""
    return float(lib().mint_flops_per_kick_pass(int(n)))


def selftest_arith(op, a, b=None, device=0):
    """Device square root of squared distances (op 0), normal-range
    quotients a / b (op 1) or the kick's square root of positive finite
    squared distances (op 2), for checks against the host's IEEE results."""
    a = np.ascontiguousarray(a, np.float64)
    out = np.empty_like(a)
    bb = np.ascontiguousarray(b, np.float64) if b is not None else None
    rc = lib().mint_selftest_arith(device, op, a.size, _d(a), _d(bb) if bb is not None else None, _d(out))
    if rc != 0:
        raise MintError(f"mint_selftest_arith failed ({rc})")
    return out


def elements_to_cartesian(mu, inc, node, argp, manom, ecc, sma):
    n = len(inc)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (inc, node, argp, manom, ecc, sma)]
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    rc = lib().mint_elements_to_cartesian(n, mu, *[_d(a) for a in arrs], _d(x), _d(v))
    if rc != 0:
        raise MintError("mint_elements_to_cartesian failed")
    return x, v


def cartesian_to_elements(mu, x, v):
    x = np.ascontiguousarray(x, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    n = len(x)
    out = [np.empty(n) for _ in range(6)]
    rc = lib().mint_cartesian_to_elements(n, mu, _d(x), _d(v), *[_d(a) for a in out])
    if rc != 0:
        raise MintError("mint_cartesian_to_elements failed")
    return tuple(out)  # inc, node, argp, mean anomaly, ecc, sma


def hill_radii(M, n_hill, m, x, v):
    m = np.ascontiguousarray(m, np.float64).ravel()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    rc = np.empty(len(m))
    lib().mint_hill_radii(len(m), M, n_hill, _d(m), _d(x), _d(v), _d(rc))
    return rc
