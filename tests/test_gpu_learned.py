"""Learned parts on the GPU: the FP32 forward/backward kernels against the
FP64 reference (oracle/nn_oracle.py), training, and the two result-neutrality
guarantees (the exact Newton tolerance and the exact distance check decide),
plus what the trained Kepler network gains over the series guess."""
import os

import numpy as np
import pytest

import mint
import mint_workloads as W
import nn_oracle as N
import oracle as O
import train_learned as T
from gpu_cases import encounter_system, run_gpu

pytestmark = pytest.mark.gpu
HAVE_WEIGHTS = os.path.exists(T.DEFAULT_OUT)


@pytest.mark.parametrize("which", [0, 1])
def test_gpu_gradients_match_fp64_reference(which):
    rng = np.random.default_rng(11 + which)
    w = rng.normal(scale=0.3, size=N.nparams(which)).astype(np.float32)
    X = rng.normal(size=(3000, 4 if which == 0 else 6)).astype(np.float32)
    t = (rng.normal(size=3000) if which == 0 else (rng.uniform(size=3000) < 0.3)).astype(np.float32)
    with mint.Mint() as g:
        g.set_nn_weights(which, w)
        loss, grad = g.nn_loss_grad(which, X, t)
        out = g.nn_forward(which, X[:100])
    ref = N.grad(which, w.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(N.loss(which, w, X, t), rel=1e-4)
    assert np.max(np.abs(grad - ref)) <= 2e-4 * np.max(np.abs(ref))
    assert np.allclose(out, N.forward(which, w, X[:100]), rtol=1e-4, atol=1e-5)


def test_training_reduces_loss():
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(50000, 4)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) * X[:, 1] + 0.5 * X[:, 2] ** 2 - X[:, 3]).astype(np.float32)
    with mint.Mint() as g:
        losses = g.train(mint.NET_KEPLER, X, t, epochs=8, batch=1024, lr=3e-3, seed=5)
    assert losses[-1] < 0.1 * losses[0], losses


def _weights(g):
    if not HAVE_WEIGHTS:
        pytest.skip("no trained weights (run p9-integrator_amd/train_learned.py on the GPU)")
    return T.apply(g)


def test_learned_guess_is_bitwise_neutral_on_the_sweep():
    st = W.p9_sweep(systems=np.arange(12), tp_scale=1 / 8)
    outs = []
    for guess in (mint.GUESS_SERIES, mint.GUESS_LEARNED):
        with mint.Mint(dt=0.5, kepler_guess=guess, chunk_steps=32, block_size=128, tps_per_lane=1) as g:
            _weights(g)
            g.upload(st)
            g.step(100)
            outs.append(g.download())
    assert np.array_equal(outs[0].tx, outs[1].tx) and np.array_equal(outs[0].tv, outs[1].tv)


def test_learned_guess_agrees_where_it_engages_and_saves_newton_iterations():
    st = W.inner_belt(4000)
    res = {}
    for guess in (mint.GUESS_SERIES, mint.GUESS_LEARNED):
        with mint.Mint(dt=0.5, kepler_guess=guess, collect_stats=1, chunk_steps=50, block_size=128,
                       tps_per_lane=1) as g:
            _weights(g)
            g.upload(st)
            g.step(100)
            res[guess] = (g.download(), g.stats())
    a, sa = res[mint.GUESS_SERIES]
    b, sb = res[mint.GUESS_LEARNED]
    ref, _, _ = O.run(st, 100, 0.5, guess=1)
    rel = np.max(np.abs(b.tx - ref["tx"]) / np.abs(ref["tx"]).max(axis=1, keepdims=True))
    assert rel <= 1e-12, rel
    assert np.array_equal(a.tx, ref["tx"])
    it_series = sa["newton_iters"] / sa["kepler_solves"]
    it_learned = sb["newton_iters"] / sb["kepler_solves"]
    print(f"mean Newton iterations per drift: series guess {it_series:.4f}, learned guess {it_learned:.4f}")
    assert it_learned < it_series


def test_scorer_prebucketing_is_bitwise_neutral():
    st = encounter_system(n_tp=400, seed=9)
    outs = []
    stats = []
    for scorer in (0, 1):
        with mint.Mint(dt=0.5, use_scorer=scorer, collect_stats=1, chunk_steps=20, block_size=64,
                       tps_per_lane=1) as g:
            _weights(g)
            g.upload(st)
            g.step(60)
            outs.append(g.download())
            stats.append(g.stats())
    assert stats[1]["prebucketed"] > 0
    assert np.array_equal(outs[0].tx, outs[1].tx) and np.array_equal(outs[0].tv, outs[1].tv)


@pytest.mark.parametrize("nm", [1, 2, 3, 4, 6, 7, 8])
def test_scorer_at_every_massive_count_matches_the_oracle(nm):
    """The trained scorer's queue head at 1..8 massive bodies (its pair loop
    and the pre-bucket kernel are instantiated per count): TPs around the
    first and last planet are picked, and every TP and body still ends bit
    for bit where the oracle puts it."""
    from gpu_cases import planet_cloud

    st = planet_cloud(nm, n_tp=128, seed=3)
    print(f"scorer, planet cloud nm={nm}", flush=True)
    with mint.Mint(dt=0.5, use_scorer=1, collect_stats=1, chunk_steps=16, block_size=64, tps_per_lane=1) as g:
        _weights(g)
        g.upload(st)
        g.step(40)
        out = g.download()
        s = g.stats()
        status = g.status()
    ref, rst, rs = O.run(st, 40, 0.5, guess=1)
    assert s["prebucketed"] > 0
    assert s["bs_steps"] == rs["bs_steps"]
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.mx, ref["mx"]) and np.all(status == rst)


def test_scorer_moves_encounters_out_of_the_main_kernel():
    """What the trained scorer gains: TPs about to meet a massive body start
    the window in the encounter queue instead of being deferred mid-window
    (which idles their lane in the main kernel for the rest of the window).
    On a slice of the sweep at the trained window length, mid-window
    deferrals fall to a small fraction, with results bitwise unchanged."""
    st = W.p9_sweep(systems=np.arange(0, 4096, 16), tp_scale=1 / 2)
    res = []
    for scorer in (0, 1):
        with mint.Mint(dt=0.5, use_scorer=scorer, collect_stats=1, chunk_steps=T.window(), block_size=256,
                       tps_per_lane=1) as g:
            _weights(g)
            g.upload(st)
            g.step(4 * T.window())
            res.append((g.download(), g.stats()))
    (o0, s0), (o1, s1) = res
    print(f"deferred mid-window {s0['deferred']} -> {s1['deferred']}, queue-head picks {s1['prebucketed']}",
          flush=True)
    assert s0["deferred"] > 30
    assert s1["deferred"] <= 0.2 * s0["deferred"]
    assert np.array_equal(o0.tx, o1.tx) and np.array_equal(o0.tv, o1.tv)


def test_scorer_far_pair_bound_covers_the_network_and_skipping_changes_no_pick():
    """Far pairs are left out of the scorer's work only under a proof: the
    library's interval bound of the logit over every pair with both distance
    features at or above its cutoff (checked here against 2e5 sampled pairs
    of that box, corners included, on the FP64 reference network) lies below
    the threshold.  With the trained weights the skip is on, and on a sweep
    slice it picks exactly the TPs the full scoring picks, with bitwise equal
    results."""
    st = W.p9_sweep(systems=np.arange(0, 4096, 32), tp_scale=1 / 2)
    res = []
    for mode in (0, 1):
        with mint.Mint(dt=0.5, use_scorer=1, collect_stats=1, chunk_steps=T.window(), block_size=256,
                       tps_per_lane=1) as g:
            thr = _weights(g)
            bound, cut, active = g.scorer_skip(mode)
            w = g.nn_weights(mint.NET_SCORER).astype(np.float64)
            g.upload(st)
            g.step(2 * T.window())
            res.append((g.download(), g.stats(), bound, cut, active))
    (o0, s0, b0, c0, a0), (o1, s1, b1, c1, a1) = res
    assert not a0 and a1 and b0 == b1 and c0 == c1 <= 8.0 and b1 < thr
    rng = np.random.default_rng(4)
    lo = np.array([c1, -4.0, 0.0, c1, 0.0, 0.0])
    hi = np.array([8.0, 4.0, 4.0, 8.0, 4.0, 2.0])
    X = rng.uniform(lo, hi, size=(200000, 6))
    for k in range(6):  # faces of the box too
        X[k::13, k] = rng.choice([lo[k], hi[k]], size=X[k::13].shape[0])
    assert np.max(N.forward(1, w, X)) < b1
    print(f"far-pair cutoff {c1} (pairs beyond {8 * c1} r_c), bound {b1:.3f} < threshold {thr:.3f}; prebucketed "
          f"{s0['prebucketed']} (all pairs) vs {s1['prebucketed']} (far pairs left out)", flush=True)
    assert s0["prebucketed"] == s1["prebucketed"] > 0
    assert np.array_equal(o0.tx, o1.tx) and np.array_equal(o0.tv, o1.tv)


def test_learned_guess_on_the_sweep_has_no_newton_iteration_to_save():
    """The Kepler network on its own target workload (the 4096-system sweep,
    learned+batched form): every drift there, test particles and giants
    alike, already converges in the first Newton evaluation from the
    third-order series guess, so no warm start can take iterations away.
    Measured on a sweep slice over two launch windows: the plain form's crude
    guess needs about 1.75 evaluations per TP drift, the series and the
    learned guesses exactly 1 (the learned guess's gate, tau > 0.02, never
    opens for a TP at r >= 30 AU, so the results are bitwise those of the
    series guess), and the oracle counts exactly one evaluation for every
    massive body's drift too (the same algorithm, bit for bit, as the
    ephemeris kernel's)."""
    st = W.p9_sweep(systems=np.arange(0, 4096, 64), tp_scale=1 / 4)
    steps = 64
    res = {}
    for name, form, guess in (("crude", mint.FORM_FLAT, mint.GUESS_CRUDE),
                              ("series", mint.FORM_BATCHED, mint.GUESS_SERIES),
                              ("learned", mint.FORM_BATCHED, mint.GUESS_LEARNED)):
        with mint.Mint(dt=0.5, form=form, kepler_guess=guess, collect_stats=1, chunk_steps=32, block_size=256,
                       tps_per_lane=1) as g:
            if guess == mint.GUESS_LEARNED:
                _weights(g)
            g.upload(st)
            g.step(steps)
            s = g.stats()
            res[name] = (s["newton_iters"] / s["kepler_solves"], g.download())
    giants = mint.State(np.zeros(st.n_systems + 1, np.int64), st.m, st.mx, st.mv, np.zeros((0, 3)),
                        np.zeros((0, 3)))
    _, _, os_ = O.run(giants, steps, 0.5)
    mass = os_["massive_newton_iters"] / os_["massive_kepler_solves"]
    print(f"Newton evaluations per drift on the sweep: TPs crude {res['crude'][0]:.4f}, series "
          f"{res['series'][0]:.4f}, learned {res['learned'][0]:.4f}; massive bodies (oracle) {mass:.4f}", flush=True)
    assert res["crude"][0] > 1.5
    assert res["series"][0] == 1.0 and res["learned"][0] == 1.0 and mass == 1.0
    a, b = res["series"][1], res["learned"][1]
    assert np.array_equal(a.tx, b.tx) and np.array_equal(a.tv, b.tv) and np.array_equal(a.mx, b.mx)
