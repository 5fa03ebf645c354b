"""Launch-shape autotuner of the batched form: the choice is a valid shape,
held parameters stay held, the cache file returns the same choice, the device
state is unchanged by tuning, and results do not depend on the shape chosen."""
import numpy as np
import pytest

import mint
import mint_workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_choice_cache_and_invariance(tmp_path):
    st = W.p9_sweep(systems=np.arange(6), tp_scale=1 / 8)
    cache = str(tmp_path / "tune.cache")
    print("autotune: 6 systems, block held at 128", flush=True)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED) as g:
        g.upload(st)
        base = g.download()                       # the upload/download round trip alone
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, block_size=128) as g:
        g.upload(st)
        best = g.autotune(cache)
        assert best[0] == 128 and best[1] in (1, 2, 4) and best[2] in (32, 64, 128), best
        after = g.download()
        assert np.array_equal(after.tx, base.tx) and np.array_equal(after.tv, base.tv)
        assert np.array_equal(after.mx, base.mx) and np.array_equal(after.mv, base.mv)
        assert g.autotune(cache) == best          # read back from the cache file
        g.step(100)
        tuned = g.download()
    other = (1, 32) if (best[1], best[2]) != (1, 32) else (2, 64)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, block_size=64, tps_per_lane=other[0],
                   chunk_steps=other[1]) as g:
        g.upload(st)
        g.step(100)
        ref = g.download()
    assert np.array_equal(tuned.tx, ref.tx) and np.array_equal(tuned.tv, ref.tv)
    assert np.array_equal(tuned.mx, ref.mx) and np.array_equal(tuned.mv, ref.mv)


def test_autotune_one_window_shape_and_launch_shape(tmp_path):
    """With block and TPs per lane free, the tuner also picks them for calls
    of one launch window: a call of at most one window reports (and uses)
    that shape, a longer call the other; the cache file holds both, and a
    call of one window gives the same bits as the same call at a held
    shape."""
    st = W.p9_sweep(systems=np.arange(6), tp_scale=1 / 8)
    cache = str(tmp_path / "tune.cache")
    print("autotune: 6 systems, window held at 32, one-window shape", flush=True)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=32) as g:
        g.upload(st)
        best = g.autotune(cache)
        one = g.launch_shape(20)
        assert best[2] == 32 and g.launch_shape(100) == best
        assert one[0] in (64, 128, 256, 512) and one[1] in (1, 2, 4) and one[2] == 32, one
        g.step(20)
        out = g.download()
    fields = open(cache).read().split()
    assert [int(x) for x in fields[-5:]] == [best[0], best[1], best[2], one[0], one[1]]
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=32) as g:
        g.upload(st)
        assert g.autotune(cache) == best and g.launch_shape(20) == one   # read back
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, block_size=64 if one[0] != 64 else 128,
                   tps_per_lane=2 if one[1] != 2 else 1, chunk_steps=32) as g:
        g.upload(st)
        g.step(20)
        ref = g.download()
    assert np.array_equal(out.tx, ref.tx) and np.array_equal(out.tv, ref.tv)
    assert np.array_equal(out.mx, ref.mx) and np.array_equal(out.mv, ref.mv)
