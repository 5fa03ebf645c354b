"""The learned Kepler guess's gate, dt^2 mu / r0^3 > KM_TAU_GATE^2 evaluated as
(dt*dt) * (mu * (ir0 * (ir0 * ir0))) in doubles, is preceded on the device by
the comparison 1/r0 > gate_ir0 with gate_ir0 = cbrt(KM_TAU_GATE^2 / (mu dt^2))
(1 - 1e-6) formed on the host.  That comparison must never shut a gate that
would open: at and below the bound the double-rounded product stays below the
threshold, and just above the exact threshold the bound lets the gate open."""
import numpy as np

KM_TAU_GATE = 0.02
G2 = KM_TAU_GATE * KM_TAU_GATE


def _bound(mu, dt):
    gd = mu * (dt * dt)
    return np.cbrt(G2 / gd) * (1.0 - 1e-6) if gd > 0.0 else np.inf


def _gate(ir0, mu, dt):
    return (dt * dt) * (mu * (ir0 * (ir0 * ir0))) > G2


def test_gate_pretest_never_shuts_an_open_gate():
    rng = np.random.default_rng(7)
    mus = [1.0, 4.0 * np.pi ** 2, 39.476926421373, 2.9591220828559115e-4, 1e-3, 7.5]
    dts = [0.5, 0.8219, -0.5, 1e-3, 10.0, 0.01, 365.25]
    for mu in mus:
        for dt in dts:
            b = _bound(mu, dt)
            exact = np.cbrt(G2 / (mu * dt * dt))
            # at and just below the bound (including its float neighbours)
            below = np.concatenate([
                b * (1.0 - rng.uniform(0.0, 1e-3, 20000)),
                [b, np.nextafter(b, 0.0), np.nextafter(np.nextafter(b, 0.0), 0.0)],
            ])
            assert not _gate(below, mu, dt).any(), (mu, dt)
            # the bound sits within 2e-6 below the exact threshold, so every
            # 1/r0 that opens the gate passes the comparison
            assert b < exact and b > exact * (1.0 - 2e-6)
            above = exact * (1.0 + rng.uniform(1e-9, 1e-3, 20000))
            opened = _gate(above, mu, dt)
            assert opened.all(), (mu, dt)
            assert (above[opened] > b).all()


def test_gate_pretest_degenerate_steps():
    # dt = 0 (or mu = 0): the bound is infinite and the gate itself stays shut
    assert _bound(1.0, 0.0) == np.inf
    assert not _gate(np.array([1e10]), 1.0, 0.0).any()
