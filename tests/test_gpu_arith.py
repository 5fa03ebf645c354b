"""The device's square root of squared distances and its normal-range
quotient (the shortened forms of the compiler's correctly rounded sequences
used on the step path) against the host's IEEE sqrt and division, bit for
bit, over the ranges the step path can reach and well beyond."""
import numpy as np
import pytest

import mint

pytestmark = pytest.mark.gpu


def _loguniform(rng, lo, hi, n):
    return np.exp2(rng.uniform(lo, hi, n))


def test_sqrt_of_squared_distances_is_ieee():
    rng = np.random.default_rng(1)
    x = np.concatenate([
        _loguniform(rng, -700.0, 700.0, 1_000_000),           # far beyond any squared distance in AU
        _loguniform(rng, -40.0, 40.0, 1_000_000),             # the physical range
        np.square(rng.integers(1, 2**26, 200_000).astype(np.float64)),  # perfect squares
        [0.0, np.inf, 1.0, 2.0, 0.25, 4.0],
    ])
    sq = np.square(rng.uniform(1.0, 2.0, 200_000))           # just above / below perfect squares
    x = np.concatenate([x, np.nextafter(sq, np.inf), np.nextafter(sq, 0.0)])
    print(f"sqrt: {x.size} inputs", flush=True)
    out = mint.selftest_arith(0, x)
    ref = np.sqrt(x)
    bad = np.flatnonzero(out.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, (bad.size, x[bad[:5]], out[bad[:5]], ref[bad[:5]])


def test_normal_range_quotient_is_ieee():
    rng = np.random.default_rng(2)
    n = 1_000_000
    a = _loguniform(rng, -400.0, 400.0, n) * rng.choice([-1.0, 1.0], n)
    b = _loguniform(rng, -400.0, 400.0, n) * rng.choice([-1.0, 1.0], n)
    a2 = _loguniform(rng, -30.0, 30.0, n)                      # the physical range
    b2 = _loguniform(rng, -30.0, 30.0, n)
    k = rng.integers(1, 2**20, 200_000).astype(np.float64)     # exact quotients
    b3 = rng.uniform(1.0, 2.0, 200_000)
    A = np.concatenate([a, a2, k * b3, np.ones(5)])
    B = np.concatenate([b, b2, b3, [1.0, 3.0, 7.0, 0.1, 1e-300 * 1e280]])
    print(f"division: {A.size} inputs", flush=True)
    out = mint.selftest_arith(1, A, B)
    ref = A / B
    bad = np.flatnonzero(out.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, (bad.size, A[bad[:5]], B[bad[:5]], out[bad[:5]], ref[bad[:5]])


def test_kick_square_root_and_its_special_cases():
    """The kick's square root (no zero / infinity cases) is IEEE on every
    positive finite squared distance; at zero and infinity it gives NaN, and
    so does the reciprocal of a zero or infinite product of denominators, so
    the kick's NaN weights there do not depend on which root was taken."""
    rng = np.random.default_rng(3)
    x = np.concatenate([
        _loguniform(rng, -700.0, 700.0, 500_000),
        _loguniform(rng, -40.0, 40.0, 500_000),
        np.square(rng.integers(1, 2**26, 100_000).astype(np.float64)),
    ])
    out = mint.selftest_arith(2, x)
    ref = np.sqrt(x)
    bad = np.flatnonzero(out.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, (bad.size, x[bad[:5]], out[bad[:5]], ref[bad[:5]])
    special = np.array([0.0, np.inf])
    assert np.isnan(mint.selftest_arith(2, special)).all()
    assert (mint.selftest_arith(0, special) == special).all()
    assert np.isnan(mint.selftest_arith(1, np.ones(3), np.array([0.0, np.inf, np.nan]))).all()
