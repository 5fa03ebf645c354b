"""The kick's far-field shortcut: a TP with r^2 above r_c^2 (1 + 1e-9) (the
parameter block's far radius) is given K = 1 without forming
y = (r - 0.1 r_c) / (0.9 r_c).  That matches the oracle, which always forms
y, only if the computed y is never below 1 there.  Checked with the same
floating-point operations, at the smallest r^2 above the far radius, over
r_c from 1e-6 to 1e6 AU (the margin is ~5e-10 against rounding of ~1e-16)."""
import numpy as np


def test_far_radius_implies_full_weight():
    rng = np.random.default_rng(3)
    R = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 2_000_000))
    rclo = 0.1 * R
    rcwinv = 1.0 / (0.9 * R)
    far2 = (R * R) * (1.0 + 1e-9)
    r2 = np.nextafter(far2, np.inf)          # the nearest a far TP can be
    r = np.sqrt(r2)
    y = (r - rclo) * rcwinv
    assert np.all(y >= 1.0)
    assert np.min(y) > 1.0 + 4e-10
