"""Edge cases of the batched form against the oracle, bit for bit: step
counts that are not a multiple of the window (and shorter than one), one-step
windows, a single TP, and systems without TPs between systems with them,
with close encounters carried across window boundaries."""
import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import encounter_system, run_gpu

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
