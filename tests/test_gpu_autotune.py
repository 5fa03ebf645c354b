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
