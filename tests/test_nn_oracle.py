"""The FP64 network reference: analytic gradients agree with central finite
differences (so the GPU gradient check below compares against a pinned
reference)."""
import numpy as np
import pytest

import nn_oracle as N


@pytest.mark.parametrize("which", [0, 1])
def test_analytic_gradient_matches_finite_differences(which):
    rng = np.random.default_rng(3)
    w = rng.normal(scale=0.3, size=N.nparams(which))
    X = rng.normal(size=(64, 4 if which == 0 else 6))
    t = rng.normal(size=64) if which == 0 else (rng.uniform(size=64) < 0.3).astype(float)
    g = N.grad(which, w, X, t)
    idx = rng.choice(len(w), 40, replace=False)
    fd = N.fd_grad(which, w, X, t, idx)
    assert np.allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


def test_layouts_have_the_library_sizes():
    assert N.nparams(0) == 4 * 32 + 32 + 32 * 32 + 32 + 32 + 1
    assert N.nparams(1) == 6 * 16 + 16 + 16 + 1
