"""mint_step_noalloc, the allocation-free variant of mint_step: the flat
per-step schedule with the library's fixed device scratch.  Its results equal
mint_step's with the flat form bit for bit, on two calls in a row (the second
on a smaller batch uploaded into the same context), and a batch larger than
the scratch is refused with MINT_ELIMIT."""
import numpy as np
import pytest

import mint
import mint_workloads as W

pytestmark = pytest.mark.gpu


def _flat(st, nsteps, guess):
    with mint.Mint(dt=0.5, form=mint.FORM_FLAT, kepler_guess=guess) as g:
        g.upload(st)
        g.step(nsteps)
        return g.download(), g.status()


def _same(a, b):
    assert np.array_equal(a.tx, b.tx) and np.array_equal(a.tv, b.tv)
    assert np.array_equal(a.mx, b.mx) and np.array_equal(a.mv, b.mv)


@pytest.mark.parametrize("guess", [mint.GUESS_CRUDE, mint.GUESS_SERIES])
def test_noalloc_step_equals_the_flat_form_on_two_calls_the_second_smaller(guess):
    big = W.p9_sweep(systems=np.arange(8), tp_scale=1 / 8)
    small = W.p9_sweep(systems=np.arange(20, 23), tp_scale=1 / 16)
    print(f"noalloc: guess {guess}, {big.n_tp} TPs x 30 steps, then {small.n_tp} TPs x 20 steps", flush=True)
    ref_big, st_big = _flat(big, 30, guess)
    ref_small, st_small = _flat(small, 20, guess)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, kepler_guess=guess) as g:
        g.upload(big)
        g.step_noalloc(30)
        out_big, s_big = g.download(), g.status()
        g.upload(small)
        g.step_noalloc(20)
        out_small, s_small = g.download(), g.status()
    _same(out_big, ref_big)
    _same(out_small, ref_small)
    assert np.array_equal(s_big, st_big) and np.array_equal(s_small, st_small)


def test_noalloc_step_refuses_a_batch_beyond_its_scratch():
    # one massive body and no TPs: 3 + 16 doubles of scratch per system, so
    # 60,000 systems need more than 2^20
    n = 60000
    rng = np.random.default_rng(1)
    a = rng.uniform(1.0, 2.0, n)
    mx = np.stack([a, np.zeros(n), np.zeros(n)], 1)[:, None, :]
    mv = np.stack([np.zeros(n), 1.0 / np.sqrt(a), np.zeros(n)], 1)[:, None, :]
    st = mint.State(np.zeros(n + 1, np.int64), np.full((n, 1), 1e-6), mx, mv, np.zeros((0, 3)), np.zeros((0, 3)))
    with mint.Mint(dt=0.5) as g:
        g.upload(st)
        with pytest.raises(mint.MintError, match="scratch") as ei:
            g.step_noalloc(1)
        assert "(-6)" in str(ei.value)
