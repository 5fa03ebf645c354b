"""Edge cases of the batched form against the oracle, bit for bit: step
counts that are not a multiple of the window (and shorter than one), one-step
windows, a single TP, and systems without TPs between systems with them,
with close encounters carried across window boundaries; non-finite TPs; TPs
on unbound and near-parabolic orbits, and on close-in orbits whose drifts
take the Stumpff functions' argument quartering (plain form too)."""
import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import encounter_system, run_gpu, unbound_cloud

pytestmark = pytest.mark.gpu


def _check(st, nsteps, **kw):
    out, status, s = run_gpu(st, nsteps, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, **kw)
    ref, rst, rs = O.run(st, nsteps, 0.5, guess=1)
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.mx, ref["mx"]) and np.array_equal(out.mv, ref["mv"])
    assert np.all(status == rst)
    assert s["bs_steps"] == rs["bs_steps"]
    return rs


@pytest.mark.parametrize("nsteps,window", [(37, 16), (5, 16), (9, 1), (33, 32)])
def test_window_lengths(nsteps, window):
    st = encounter_system(n_tp=128, seed=21)
    print(f"encounter cloud, {nsteps} steps, window {window}", flush=True)
    rs = _check(st, nsteps, chunk_steps=window, block_size=64, tps_per_lane=1)
    assert rs["bs_steps"] > 0


def test_single_tp_and_empty_systems():
    base = W.p9_sweep(systems=np.arange(5), tp_scale=1 / 16)
    cloud = encounter_system(n_tp=64, seed=4)
    # systems: 0 with one TP, 1 without TPs, 2 the encounter cloud, 3 without TPs, 4 a sweep slice
    parts = [(base.subset([0]), slice(0, 1)), (base.subset([1]), slice(0, 0)), (cloud, slice(None)),
             (base.subset([3]), slice(0, 0)), (base.subset([4]), slice(None))]
    m, mx, mv, tx, tv, counts = [], [], [], [], [], []
    for s, sl in parts:
        m.append(s.m[0]); mx.append(s.mx[0]); mv.append(s.mv[0])
        tx.append(s.tx[sl]); tv.append(s.tv[sl]); counts.append(len(s.tx[sl]))
    st = mint.State(np.concatenate([[0], np.cumsum(counts)]).astype(np.int64), np.array(m), np.array(mx),
                    np.array(mv), np.concatenate(tx), np.concatenate(tv))
    print(f"mixed batch: TPs per system {counts}", flush=True)
    rs = _check(st, 50, chunk_steps=16, block_size=64, tps_per_lane=2)
    assert rs["bs_steps"] > 0


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_BATCHED])
def test_non_finite_tps_are_flagged_and_isolated(form):
    """A TP uploaded with a NaN position and one with an infinite velocity:
    their systems are flagged non-finite (as the oracle flags them), and every
    other TP and every massive body ends bit for bit where it ends without
    them (the encounter check never fires on a NaN or infinite separation, and
    every loop a non-finite state can reach is bounded)."""
    st = W.p9_sweep(systems=[0, 1, 2], tp_scale=1 / 32)
    bad = st.copy()
    i_nan = int(st.tp_offset[0]) + 3
    i_inf = int(st.tp_offset[2]) + 5
    bad.tx[i_nan, 0] = np.nan
    bad.tv[i_inf, 1] = np.inf
    kw = dict(chunk_steps=16, block_size=64, tps_per_lane=1) if form == mint.FORM_BATCHED else {}
    print(f"non-finite TPs {i_nan}, {i_inf}; form {form}", flush=True)
    out_c, stat_c, _ = run_gpu(st, 40, 0.5, form, mint.GUESS_SERIES, **kw)
    out_b, stat_b, _ = run_gpu(bad, 40, 0.5, form, mint.GUESS_SERIES, **kw)
    keep = np.ones(st.n_tp, bool)
    keep[[i_nan, i_inf]] = False
    assert np.array_equal(out_b.tx[keep], out_c.tx[keep]) and np.array_equal(out_b.tv[keep], out_c.tv[keep])
    assert np.array_equal(out_b.mx, out_c.mx) and np.array_equal(out_b.mv, out_c.mv)
    assert not np.any(stat_c & mint.ST_NAN)
    assert stat_b[0] & mint.ST_NAN and stat_b[2] & mint.ST_NAN and not stat_b[1] & mint.ST_NAN
    assert not np.all(np.isfinite(out_b.tx[i_nan])) and not np.all(np.isfinite(out_b.tv[i_inf]))
    _, rst, _ = O.run(bad, 40, 0.5, guess=1)
    assert np.array_equal(rst & mint.ST_NAN, stat_b & mint.ST_NAN)


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_BATCHED])
def test_unbound_and_near_parabolic_tps_match_the_oracle(form):
    """TPs on hyperbolic orbits, within 1e-9 of the parabolic limit, around
    escape and slow and nearly radial: the universal-variable drift's
    negative-alpha Stumpff branches (at 20-200 AU every |z| stays below the
    series band's 0.1; the close-in test below takes the quartering) over 100
    steps, bit for bit against the oracle."""
    st = unbound_cloud()
    print(f"unbound cloud form={form}", flush=True)
    kw = dict(chunk_steps=32, block_size=64, tps_per_lane=1) if form == mint.FORM_BATCHED else {}
    out, status, s = run_gpu(st, 100, 0.5, form, mint.GUESS_SERIES, **kw)
    ref, rst, rs = O.run(st, 100, 0.5, guess=1)
    assert np.all(np.isfinite(ref["tx"])) and np.abs(ref["tx"]).max() > 200.0
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.mx, ref["mx"]) and np.all(status == rst)
    assert s["newton_iters"] == rs["newton_iters"] > rs["kepler_solves"]


def _close_in(n=256, seed=23):
    """The giants of sweep system 0 with TPs on close-in orbits (a 0.05-1.2
    AU, e < 0.3, i < 10 deg): at dt = 0.5 yr a drift spans up to several
    orbits, so z = beta s^2 of the Kepler solve runs from below the series
    band's 0.1 to beyond 1e3 (up to six quarterings of the argument)."""
    st = W.p9_sweep(systems=[0], tp_scale=1 / 64).copy()
    rng = np.random.default_rng([20261203, 11, seed])
    a = np.exp(rng.uniform(np.log(0.05), np.log(1.2), n))
    e = rng.uniform(0.0, 0.3, n)
    inc = np.radians(rng.uniform(0.0, 10.0, n))
    ang = rng.uniform(0.0, 2 * np.pi, (3, n))
    st.tx, st.tv = mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], e, a)
    st.tp_offset = np.array([0, n], np.int64)
    return st


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_BATCHED])
def test_close_in_orbits_take_the_stumpff_quartering_and_match_the_oracle(form):
    st = _close_in()
    # the first drift of every TP, from the oracle's converged root: z = beta s^2
    z = []
    for x, v in zip(st.tx, st.tv):
        _, _, _, s = O.kepler_drift(x, v, 1.0, 0.5)
        z.append((2.0 / np.linalg.norm(x) - v @ v) * s * s)
    z = np.abs(np.array(z))
    print(f"close-in cloud form={form}: |z| from {z.min():.3g} to {z.max():.3g}, "
          f"{np.mean(z > 0.1):.2f} of the drifts above 0.1", flush=True)
    assert np.mean(z > 0.1) > 0.5 and np.max(z) > 0.1 * 4 ** 5
    kw = dict(chunk_steps=16, block_size=64, tps_per_lane=1) if form == mint.FORM_BATCHED else {}
    out, status, s = run_gpu(st, 50, 0.5, form, mint.GUESS_SERIES, **kw)
    ref, rst, rs = O.run(st, 50, 0.5, guess=1)
    assert np.all(np.isfinite(ref["tx"])) and not np.any(rst & mint.ST_NEWTON)
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.mx, ref["mx"]) and np.all(status == rst)
    assert s["newton_iters"] == rs["newton_iters"] > rs["kepler_solves"]
