"""The far radius r_c^2 (1 + 1e-3) of the TP kick.  Beyond it the kick gives
K = 1 without forming y = (r - 0.1 r_c) / (0.9 r_c), and the next step's
encounter check is skipped.  Both match the oracle (which forms y and checks
every step) only if, with the same floating-point operations, y is never
below 1 there and the check can never fire; both are checked here at the
nearest a TP beyond the far radius can be."""
import numpy as np


def test_far_radius_implies_full_weight():
    rng = np.random.default_rng(3)
    R = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 2_000_000))
    rclo = 0.1 * R
    rcwinv = 1.0 / (0.9 * R)
    far2 = (R * R) * (1.0 + 1e-3)
    r2 = np.nextafter(far2, np.inf)          # the nearest a far TP can be
    r = np.sqrt(r2)
    y = (r - rclo) * rcwinv
    assert np.all(y >= 1.0)
    assert np.min(y) > 1.0 + 4e-4


def _fma(a, b, c):
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _dot3(a, b):
    return _fma(a[0], b[0], _fma(a[1], b[1], a[2] * b[2]))


def test_far_kick_rules_out_the_next_encounter_check():
    """The main kernel skips a step's encounter check when the previous kick
    found every pair beyond the far radius.  The check measures the
    separation of x' = fma(h/2, Q1, X) and Y = fma(h/2, Q1, XE) (the same jump
    applied to the TP and the body); the kick measured XE - X.  Emulated with
    exactly rounded fma, the check never fires just beyond the far radius,
    for bodies up to 1e6 AU out, r_c down to 2e-4 AU and jumps up to 1 AU."""
    rng = np.random.default_rng(5)
    n_tested = 0
    for _ in range(4000):
        rc = float(np.exp(rng.uniform(np.log(2e-4), np.log(100.0))))
        rc2 = rc * rc
        far2 = rc2 * (1.0 + 1e-3)
        u = rng.normal(size=3)
        XE = u / np.linalg.norm(u) * float(np.exp(rng.uniform(np.log(0.1), np.log(1e6))))
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        X = XE - w * rc * np.sqrt(1.0 + 1e-3) * (1.0 + rng.uniform(0.0, 1e-9))
        d = XE - X
        if not _dot3(d, d) > far2:
            continue
        n_tested += 1
        hh = 0.25
        Q = rng.normal(size=3) * float(np.exp(rng.uniform(np.log(1e-6), np.log(4.0))))
        xp = np.array([_fma(hh, Q[c], X[c]) for c in range(3)])
        Y = np.array([_fma(hh, Q[c], XE[c]) for c in range(3)])
        dc = Y - xp
        assert not _dot3(dc, dc) < rc2, (rc, XE, X, Q)
    assert n_tested > 1000
