"""Captured launch graphs: a cached graph is replayed only while the constants
it holds by value are unchanged (a checkpoint of another step length loaded
into a context that has already stepped), a graph built ahead of its call by
mint_prepare gives the same bits, and the per-launch main-kernel timing mode
(events around each launch inside the replayed graph) changes no result and
reports one timed launch per window."""
import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import embryo_ring

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.tx, b["tx"] if isinstance(b, dict) else b.tx)
    assert np.array_equal(a.tv, b["tv"] if isinstance(b, dict) else b.tv)
    assert np.array_equal(a.mx, b["mx"] if isinstance(b, dict) else b.mx)
    assert np.array_equal(a.mv, b["mv"] if isinstance(b, dict) else b.mv)


def test_checkpoint_of_another_dt_replays_no_stale_graph(tmp_path):
    st = W.p9_sweep(systems=np.arange(3), tp_scale=1 / 8)
    kw = dict(form=mint.FORM_BATCHED, chunk_steps=16, block_size=64, tps_per_lane=1)
    path = str(tmp_path / "dt025.ckpt")
    print("checkpoint: batched form, dt 0.25 checkpoint into a dt 0.5 context", flush=True)
    with mint.Mint(dt=0.25, **kw) as b:
        b.upload(st)
        b.step(10)
        b.save_checkpoint(path, 10)
    with mint.Mint(dt=0.5, **kw) as a:
        a.upload(st)
        a.step(40)            # captures the 40-step graph with h = 0.5
        assert a.load_checkpoint(path) == 10
        a.step(40)            # same shape, same step count: must not replay h = 0.5
        out = a.download()
    ref, _, _ = O.run(st, 50, 0.25, guess=1)
    _same(out, ref)


def test_many_body_checkpoint_of_another_dt_replays_no_stale_graph(tmp_path):
    st = embryo_ring(48, seed=3)
    rad = st.meta["radius"]
    path = str(tmp_path / "emb.ckpt")
    print("checkpoint: many-body engine, dt 0.001 checkpoint into a dt 0.002 context", flush=True)
    with mint.Mint(dt=0.001, bs_tol=1e-11) as b:
        b.upload(st)
        b.set_radii(rad)
        b.step(5)
        b.save_checkpoint(path, 5)
    with mint.Mint(dt=0.002, bs_tol=1e-11) as a:
        a.upload(st)
        a.set_radii(rad)
        a.step(70)            # long enough to capture the many-body step graph
        assert a.load_checkpoint(path) == 5
        a.step(70)
        out = a.download()
    ref, _, _ = O.run(st, 75, 0.001, bs_tol=1e-11, radii=rad)
    assert np.array_equal(out.mx, ref["mx"]) and np.array_equal(out.mv, ref["mv"])


def test_prepared_graph_gives_the_same_bits():
    st = W.p9_sweep(systems=np.arange(4), tp_scale=1 / 8)
    kw = dict(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=16, block_size=64, tps_per_lane=1)
    outs = []
    for prep in (False, True):
        with mint.Mint(**kw) as g:
            g.upload(st)
            if prep:
                g.prepare(37)
            g.step(37)
            outs.append(g.download())
    _same(outs[1], outs[0])


@pytest.mark.parametrize("use_graph", [1, 0])
def test_main_kernel_timing_mode_changes_nothing_and_times_every_window(use_graph):
    st = W.p9_sweep(systems=np.arange(6), tp_scale=1 / 4)
    kw = dict(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=32, block_size=128, tps_per_lane=1, use_graph=use_graph)
    outs, tms = [], []
    for stats in (0, 4):
        with mint.Mint(collect_stats=stats, **kw) as g:
            g.upload(st)
            g.prepare(100)
            g.reset_counters()
            g.step(100)
            outs.append(g.download())
            tms.append(g.timing())
    _same(outs[1], outs[0])
    t = tms[1]
    print(f"timing mode (graph {use_graph}): {t['n_tp_main']} launches, {t['ms_tp_main']:.3f} of "
          f"{t['ms_total']:.3f} ms", flush=True)
    assert t["n_tp_main"] == 4          # windows of 32, 32, 32, 4 steps
    assert 0.0 < t["ms_tp_main"] <= t["ms_total"]
    assert tms[0]["n_tp_main"] == 0


def test_calls_of_mixed_lengths_forms_and_uploads_keep_the_queues_clear():
    """A batched call leaves the resident massive state and its first
    window's queue as the next call expects them (the call's last kernel copies
    the end state back and clears that queue): calls of one window, of
    several windows with carried deferrals, flat calls between them and a
    re-upload, at the tuned launch shapes (calls of one window and longer
    calls may each have their own, and the context keeps both tile lists), on
    an encounter-heavy system (mid-window deferrals in every window), all
    bitwise equal to the oracle over the same steps."""
    from gpu_cases import encounter_system

    st = encounter_system(n_tp=300, seed=5)
    print("queue hygiene: encounter cloud, calls 7, 40, flat 5, 16, 33 (re-upload), 3", flush=True)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=16, collect_stats=1) as g:
        g.upload(st)
        g.autotune(None)
        print(f"shapes: one window {g.launch_shape(7)}, longer {g.launch_shape(40)}", flush=True)
        g.step(7)
        g.step(40)                       # three windows, carried tails
        g.set_mode(mint.FORM_FLAT, mint.GUESS_SERIES, 0)
        g.step(5)
        g.set_mode(mint.FORM_BATCHED, mint.GUESS_SERIES, 0)
        g.step(16)
        mid = g.download()
        deferred = g.stats()["deferred"]
        g.upload(st)                     # same shape: buffers and graphs kept
        g.step(33)
        g.step(3)
        end = g.download()
    ref1, _, _ = O.run(st, 68, 0.5, guess=1)
    ref2, _, _ = O.run(st, 36, 0.5, guess=1)
    assert deferred > 0
    _same(mid, ref1)
    _same(end, ref2)
