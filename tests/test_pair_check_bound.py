"""The many-body kick kernel marks a pair slot for the exact pair check only
when r^2 lies below the system's bound: the largest r_c^2 of its bodies, or
the overlap bound (R_i + R_max)^2 with its largest physical radius R_max
(mint_mb.hip, mb_close_kernel).  The exact tests are r^2 < max(r_c,i^2,
r_c,k^2) (close_encounter_p, src/sim.cuh:589-613) and r^2 < (R_i + R_k)^2;
this checks, in the kernel's double arithmetic, that every pair passing
either exact test is marked, including pairs placed within an ulp of the
thresholds (rounding is monotonic, so the bound can only be looser)."""
import numpy as np


def _marks(r2, rc2, rad, i, k):
    rc2max = rc2.max()
    radmax = rad.max()
    so = rad[i] + radmax
    lim = np.maximum(rc2max, so * so)
    return r2 < lim


def _exact(r2, rc2, rad, i, k):
    sr = rad[i] + rad[k]
    return (r2 < np.maximum(rc2[i], rc2[k])) | (r2 < sr * sr)


def test_every_pair_inside_is_marked():
    rng = np.random.default_rng(7)
    for trial in range(200):
        n = int(rng.integers(2, 300))
        rc = rng.uniform(1e-4, 2.0, n) * 10.0 ** rng.integers(-3, 2)
        rc2 = rc * rc
        rad = np.where(rng.random(n) < 0.8, rng.uniform(0, 1e-3, n), 0.0)
        i = rng.integers(0, n, 4000)
        k = rng.integers(0, n, 4000)
        # separations near each threshold, and exactly on it, one ulp either side
        base = np.where(rng.random(4000) < 0.5, np.maximum(rc2[i], rc2[k]), (rad[i] + rad[k]) ** 2)
        r2 = base * rng.choice([1.0 - 1e-12, 1.0, 1.0 + 1e-12, 0.5, 2.0], 4000)
        r2 = np.concatenate([r2, np.nextafter(base, 0), np.nextafter(base, np.inf), base])
        ii = np.concatenate([i, i, i, i])
        kk = np.concatenate([k, k, k, k])
        inside = _exact(r2, rc2, rad, ii, kk)
        assert np.all(_marks(r2, rc2, rad, ii, kk)[inside])


def test_a_plausible_mistake_fails():
    """Planted: a bound from the smaller of the pair's radii misses pairs the
    exact test keeps, so the check above would catch such a bound."""
    rc2 = np.array([1.0, 4.0])
    rad = np.zeros(2)
    r2 = np.array([2.0])
    i, k = np.array([0]), np.array([1])
    assert _exact(r2, rc2, rad, i, k)[0]
    wrong = r2 < np.minimum(rc2[i], rc2[k])
    assert not wrong[0]
