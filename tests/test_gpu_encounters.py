"""GPU parity where the Bulirsch-Stoer close-encounter path is exercised, and
the batch-invariance properties of the batched form (bitwise)."""
import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import encounter_system, planet_cloud, run_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_FLAT, mint.FORM_BATCHED])
def test_encounter_cloud_matches_oracle(form):
    st = encounter_system()
    print(f"encounter cloud form={form}", flush=True)
    kw = dict(chunk_steps=16, block_size=64, tps_per_lane=2) if form == mint.FORM_BATCHED else {}
    out, status, s = run_gpu(st, 40, 0.5, form, mint.GUESS_SERIES, **kw)
    ref, rst, rs = O.run(st, 40, 0.5, guess=1)
    assert rs["bs_steps"] > 1000, rs               # the BS path is really taken
    assert s["bs_steps"] == rs["bs_steps"]
    assert s["bs_rhs_evals"] == rs["bs_rhs_evals"]
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.mx, ref["mx"])
    assert np.all(status == rst)


def test_system_alone_equals_system_in_batch():
    st = W.p9_sweep(systems=np.arange(24), tp_scale=1 / 8)
    out, _, _ = run_gpu(st, 64, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=32, block_size=256,
                        tps_per_lane=1)
    for s in (0, 7, 23):
        one = st.subset([s])
        o1, _, _ = run_gpu(one, 64, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=16, block_size=64,
                           tps_per_lane=4)
        a, b = st.tp_offset[s], st.tp_offset[s + 1]
        assert np.array_equal(o1.tx, out.tx[a:b]) and np.array_equal(o1.tv, out.tv[a:b])
        assert np.array_equal(o1.mx[0], out.mx[s])


def test_permuting_tps_permutes_results():
    st = W.p9_sweep(systems=[3], tp_scale=1 / 2)
    perm = np.random.default_rng(5).permutation(st.n_tp)
    sp = st.copy()
    sp.tx = st.tx[perm]
    sp.tv = st.tv[perm]
    a, _, _ = run_gpu(st, 50, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=25, block_size=128,
                      tps_per_lane=1)
    b, _, _ = run_gpu(sp, 50, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=25, block_size=128,
                      tps_per_lane=1)
    assert np.array_equal(a.tx[perm], b.tx) and np.array_equal(a.tv[perm], b.tv)


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_BATCHED])
def test_tp_trajectory_does_not_depend_on_the_other_tps(form):
    """A TP ends bit for bit in the same place whichever other TPs share its
    system (and its encounter queue, whose order the other TPs change): every
    third TP of an encounter cloud run alone against the whole cloud."""
    st = encounter_system(n_tp=400, seed=9)
    idx = np.arange(0, st.n_tp, 3)
    sub = st.copy()
    sub.tx, sub.tv = st.tx[idx], st.tv[idx]
    sub.tp_offset = np.array([0, len(idx)], np.int64)
    kw = dict(chunk_steps=16, block_size=64, tps_per_lane=1) if form == mint.FORM_BATCHED else {}
    a, _, sa = run_gpu(st, 40, 0.5, form, mint.GUESS_SERIES, **kw)
    b, _, sb = run_gpu(sub, 40, 0.5, form, mint.GUESS_SERIES, **kw)
    assert sb["bs_steps"] > 100 and sa["bs_steps"] > sb["bs_steps"]
    assert np.array_equal(a.tx[idx], b.tx) and np.array_equal(a.tv[idx], b.tv)
    assert np.array_equal(a.mx, b.mx) and np.array_equal(a.mv, b.mv)


def test_run_twice_is_deterministic_and_launch_shape_invariant():
    st = encounter_system(n_tp=300, seed=3)
    a, _, _ = run_gpu(st, 30, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=8, block_size=64,
                      tps_per_lane=1)
    b, _, _ = run_gpu(st, 30, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=8, block_size=64,
                      tps_per_lane=1)
    c, _, _ = run_gpu(st, 30, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=30, block_size=512,
                      tps_per_lane=4)
    assert np.array_equal(a.tx, b.tx) and np.array_equal(a.tx, c.tx) and np.array_equal(a.tv, c.tv)


def test_graph_replay_equals_direct_launches():
    st = W.p9_sweep(systems=np.arange(6), tp_scale=1 / 4)
    a, _, _ = run_gpu(st, 70, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=16, block_size=128,
                      tps_per_lane=1, use_graph=1)
    b, _, _ = run_gpu(st, 70, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=16, block_size=128,
                      tps_per_lane=1, use_graph=0)
    assert np.array_equal(a.tx, b.tx) and np.array_equal(a.tv, b.tv)


def test_timed_serial_launches_and_graph_reuse_match():
    """The pipelined graph, the serial launches used when every kernel is
    timed, and two calls replaying one cached graph (TPs deferred in the last
    window of a call are finished within that call, none is carried across)
    give the same bits and the same work counts."""
    st = encounter_system(n_tp=256, seed=5)
    kw = dict(chunk_steps=16, block_size=64, tps_per_lane=1)
    print("encounter cloud: graph / timed serial / two calls, 96 steps", flush=True)
    a, _, sa = run_gpu(st, 96, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, **kw)
    b, _, sb = run_gpu(st, 96, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, stats=3, **kw)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, kepler_guess=mint.GUESS_SERIES, collect_stats=1, **kw) as g:
        g.upload(st)
        g.step(48)
        g.step(48)
        c = g.download()
        sc = g.stats()
    for o in (b, c):
        assert np.array_equal(a.tx, o.tx) and np.array_equal(a.tv, o.tv) and np.array_equal(a.mx, o.mx)
    assert sa["bs_steps"] > 0 and sa["deferred"] > 0
    for key in ("tp_steps", "bs_steps", "bs_rhs_evals", "kepler_solves"):
        assert sa[key] == sb[key] == sc[key], key


@pytest.mark.parametrize("nm", [1, 2, 3, 4, 6, 7, 8])
def test_encounters_at_every_massive_count(nm):
    """BS work of TPs beside 1..8 massive bodies (the grouped encounter kernel
    splits each right-hand side over 1, 2, 4 or 8 lanes) against the oracle."""
    st = planet_cloud(nm)
    print(f"planet cloud nm={nm}", flush=True)
    out, status, s = run_gpu(st, 40, 0.5, mint.FORM_BATCHED, mint.GUESS_SERIES, chunk_steps=16, block_size=64,
                             tps_per_lane=1)
    ref, rst, rs = O.run(st, 40, 0.5, guess=1)
    assert rs["bs_steps"] > 100, rs
    assert s["bs_steps"] == rs["bs_steps"] and s["bs_rhs_evals"] == rs["bs_rhs_evals"]
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.mx, ref["mx"]) and np.all(status == rst)
