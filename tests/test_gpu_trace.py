"""The trace logger (mint_trace_kepler, mint_trace_encounter): the training
samples it writes, checked against the oracle.  Kepler: after one step, one
sample per bound body (every TP and every massive body at stride 1), with the
target c = (s*/s_series - 1)/tau^3 from the oracle's own converged solve of
the same drift (bit for bit) and the same features.  Encounter: over one
launch window, one positive sample for every (TP, body) pair whose exact
distance check fires somewhere along the TP's Kepler path (the oracle runs
that path and check), and exactly the negatives the sampling hash keeps."""
import struct

import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import encounter_system

pytestmark = pytest.mark.gpu


def _read_checkpoint(path):
    """The DH state of a mint checkpoint (the library's internal frame)."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"MINTCKP1"
    S, NM, _ = struct.unpack_from("<3i", raw, 8)
    N, _ = struct.unpack_from("<2q", raw, 20)
    o = 8 + 12 + 16 + 40
    o += 8 * (S + 1)
    arr = np.frombuffer(raw, dtype="<f8", offset=o)
    k = S * NM
    mx = arr[3 * k:6 * k].reshape(S * NM, 3)
    mv = arr[6 * k:9 * k].reshape(S * NM, 3)
    t = arr[9 * k:9 * k + 6 * N].reshape(6, N)
    return mx, mv, t[:3].T.copy(), t[3:].T.copy()


def test_kepler_trace_samples_are_the_oracles_solves(tmp_path):
    st = W.p9_sweep(systems=np.arange(6), tp_scale=1 / 8)
    kw = dict(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=16, block_size=64, tps_per_lane=1)
    nb = st.n_tp + st.n_systems * st.n_massive
    print(f"kepler trace: {st.n_tp} TPs and {st.n_systems * st.n_massive} massive bodies, one step", flush=True)
    with mint.Mint(**kw) as g:
        g.upload(st)
        feat, tgt = g.trace_kepler(1, 2 * nb)
    path = str(tmp_path / "one_step.ckpt")
    with mint.Mint(**kw) as g:
        g.upload(st)
        g.step(1)
        g.save_checkpoint(path, 1)
    mx, mv, tx, tv = _read_checkpoint(path)
    ref = [O.kepler_sample(x, v, 1.0, 0.5) for x, v in zip(np.concatenate([tx, mx]), np.concatenate([tv, mv]))]
    ref = [r for r in ref if r is not None]
    assert feat.shape == (len(ref), 4) and tgt.shape == (len(ref),) and len(ref) > 0.9 * nb
    rf = np.array([r[0] for r in ref], np.float32)
    rt = np.array([r[1] for r in ref], np.float32)

    def order(f, t):
        return np.lexsort((f[:, 3], f[:, 2], f[:, 1], t))

    a, b = order(feat, tgt), order(rf, rt)
    assert np.array_equal(tgt[a], rt[b])                  # the converged universal variable, bit for bit
    assert np.array_equal(feat[a][:, 1:], rf[b][:, 1:])   # FP64 quotients rounded to FP32
    assert np.max(np.abs(feat[a][:, 0] - rf[b][:, 0])) <= 2e-6  # log2f of the device library vs the host's


def _hash_keeps(i, q, win):
    """The trace's one-in-64 sampling of negative pairs (32-bit arithmetic)."""
    m = 0xFFFFFFFF
    h = ((i * 2654435761) & m) ^ (((q + 1) * 40503) & m) ^ ((win * 2246822519) & m)
    h ^= h >> 15
    h = (h * 2246822519) & m
    h ^= h >> 13
    return (h & 63) == 0


def _features(st, T, M=1.0):
    """Scorer features of every (TP, body) pair at the window start, in FP64
    (the device forms them in FP32 from FP64 differences)."""
    S, NM = st.m.shape
    rc = mint.hill_radii(M, 3.0, st.m.ravel(), st.mx.reshape(-1, 3), st.mv.reshape(-1, 3)).reshape(S, NM)
    sh = (st.m[:, :, None] * st.mv).sum(1) / (M + st.m.sum(1))[:, None]
    sys_of = np.repeat(np.arange(S), np.diff(st.tp_offset))
    tv = st.tv - sh[sys_of]
    out = np.empty((st.n_tp, NM, 6))
    for j in range(NM):
        X, V, r_c = st.mx[sys_of, j], st.mv[sys_of, j] - sh[sys_of], rc[sys_of, j]
        d, u = X - st.tx, V - tv
        r = np.linalg.norm(d, axis=1)
        u2, du = (u * u).sum(1), (d * u).sum(1)
        tmin = np.clip(np.where(u2 > 0, -du / np.where(u2 > 0, u2, 1), 0), 0, T)
        rmin = np.linalg.norm(d + u * tmin[:, None], axis=1)
        R, rt = np.linalg.norm(X, axis=1), np.linalg.norm(st.tx, axis=1)
        out[:, j] = np.stack([np.minimum(r / r_c, 64) * 0.125, np.clip(du / r / r_c * T, -16, 16) * 0.25,
                              np.minimum(np.sqrt(u2) * T / r_c, 16) * 0.25, np.minimum(rmin / r_c, 64) * 0.125,
                              np.minimum(r_c / R, 1) * 4, np.minimum(rt / R, 8) * 0.25], 1)
    return out


def test_encounter_trace_labels_are_the_oracles_exact_check():
    st = encounter_system(n_tp=400, seed=9)
    C = 16
    print(f"encounter trace: {st.n_tp} TPs around two bodies, one window of {C} steps", flush=True)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=C, block_size=64, tps_per_lane=1) as g:
        g.upload(st)
        feat, lab = g.trace_encounter(C, 8192)
    assert feat.shape == (len(lab), 6) and set(np.unique(lab)) <= {0.0, 1.0}
    hits = O.trace_hits(st, C, 0.5)
    NM = st.n_massive
    bits = (hits[:, None] >> np.arange(NM, dtype=np.uint32)[None, :]) & 1
    pos = np.argwhere(bits == 1)
    keep = np.array([[_hash_keeps(i, q, 0) for q in range(NM)] for i in range(st.n_tp)])
    print(f"positives {int((lab == 1).sum())} (oracle {len(pos)}), negatives {int((lab == 0).sum())} "
          f"(hash keeps {int((keep & (bits == 0)).sum())})", flush=True)
    assert len(pos) > 50
    assert int((lab == 1).sum()) == len(pos)
    assert int((lab == 0).sum()) == int((keep & (bits == 0)).sum())
    # every oracle positive has its sample: the nearest positive feature
    # vector agrees to the FP32 feature arithmetic
    ref = _features(st, 0.5 * C)[pos[:, 0], pos[:, 1]]
    got = feat[lab == 1].astype(np.float64)
    dist = np.max(np.abs(ref[:, None, :] - got[None, :, :]), axis=2)
    assert np.max(np.min(dist, axis=1)) <= 2e-4
    assert np.max(np.min(dist, axis=0)) <= 2e-4
