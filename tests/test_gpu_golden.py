"""The GPU library reproduces the golden fixtures bit for bit (all forms)."""
import os

import numpy as np
import pytest

import mint
from golden_io import all_fixtures, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [mint.FORM_FLAT, mint.FORM_BATCHED])
@pytest.mark.parametrize("path", all_fixtures(), ids=os.path.basename)
def test_gpu_reproduces_golden(path, form):
    d, st, ref = load(path)
    print(f"golden {os.path.basename(path)} form={form}", flush=True)
    with mint.Mint(dt=d["dt"], form=form, kepler_guess=d["guess"], chunk_steps=32, block_size=128,
                   tps_per_lane=2, bs_tol=d["bs_tol"]) as g:
        g.upload(st)
        if d["radii"] is not None:
            g.set_radii(d["radii"])
        g.step(d["steps"])
        out = g.download()
        status = g.status()
    assert np.array_equal(out.mx, ref["mx"]) and np.array_equal(out.mv, ref["mv"])
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])
    assert np.array_equal(out.rc, ref["rc"])
    assert np.array_equal(status, ref["status"])
