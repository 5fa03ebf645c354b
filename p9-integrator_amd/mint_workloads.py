"""Seeded synthetic workloads for the Mercurius step (G = 1, Msun, AU, yr).

The reference's generator (utils/sim_config.py:69-207) draws from an unseeded
RNG and writes one fixed JSON; these builders are seeded per system so any
subset is reproducible on its own.  Angles are radians; elements go through
the library's host conversion (mint_elements_to_cartesian).

  earth_mars()      Sun + Earth-like + Mars-like, no TPs            (dt 1e-3)
  p9_single(n_tp)   Sun + 4 giants + Planet Nine + n_tp TPs         (dt 0.5)
  p9_sweep(S)       S independent P9 systems, lognormal TP counts   (dt 0.5)
  embryos(n)        Sun + n planetary embryos, all massive          (dt 0.002)
"""
from __future__ import annotations

import numpy as np

import mint

MASTER_SEED = 20261203
EARTH_MASS = 3.00338e-06            # utils/sim_config.py:91
GIANT_M = np.array([9.55e-4, 2.86e-4, 4.37e-5, 5.15e-5])
GIANT_A = np.array([5.2, 9.55, 19.2, 30.1])
SWEEP_SIGMA = 0.75                  # lognormal sigma of TPs per system (fixed here)
SWEEP_MEDIAN = 1024
SWEEP_CLIP = (256, 8192)
AU_CM = 1.495978707e13
MSUN_G = 1.98847e33
EMBRYO_DENSITY = 3.0                # g/cm^3


def _rng(*key):
    return np.random.default_rng([MASTER_SEED, *key])


def _tps(rng, n):
    """TPs: a ~ U[150,1000], perihelion q ~ U[30,50], i ~ U[0,40 deg], angles U[0,2pi)."""
    a = rng.uniform(150.0, 1000.0, n)
    q = rng.uniform(30.0, 50.0, n)
    e = 1.0 - q / a
    inc = np.radians(rng.uniform(0.0, 40.0, n))
    ang = rng.uniform(0.0, 2 * np.pi, (3, n))
    return mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], e, a)


def _giants(rng):
    e = rng.uniform(0.0, 0.06, 4)
    inc = np.radians(rng.uniform(0.0, 2.5, 4))
    ang = rng.uniform(0.0, 2 * np.pi, (3, 4))
    return GIANT_M.copy(), (inc, ang[0], ang[1], ang[2], e, GIANT_A.copy())


def _massive_state(m, els):
    x, v = mint.elements_to_cartesian(1.0, *els)
    return m, x, v


def earth_mars(seed_key=0) -> mint.State:
    rng = _rng(0, seed_key)
    m = np.array([3.0e-6, 3.2e-7])
    inc = np.radians(np.array([0.0, 1.85]))
    ang = rng.uniform(0.0, 2 * np.pi, (3, 2))
    x, v = mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], np.array([0.017, 0.093]),
                                      np.array([1.0, 1.524]))
    return mint.State(np.array([0, 0], np.int64), m[None], x[None], v[None], np.zeros((0, 3)), np.zeros((0, 3)),
                      meta=dict(workload="earth_mars", dt=1e-3))


def p9_single(n_tp=100_000, seed_key=0) -> mint.State:
    """Sun + giants + P9 (m 1.5e-5..3e-5, a 400..800, e 0.2..0.5, i 15..25 deg)."""
    rng = _rng(1, seed_key)
    gm, gel = _giants(rng)
    p9m = rng.uniform(1.5e-5, 3e-5)
    p9 = (np.radians([rng.uniform(15, 25)]), *rng.uniform(0, 2 * np.pi, (3, 1)), [rng.uniform(0.2, 0.5)],
          [rng.uniform(400, 800)])
    m = np.concatenate([[p9m], gm])
    els = [np.concatenate([np.asarray(p, float), g]) for p, g in zip(p9, gel)]
    m, x, v = _massive_state(m, els)
    tx, tv = _tps(rng, n_tp)
    return mint.State(np.array([0, n_tp], np.int64), m[None], x[None], v[None], tx, tv,
                      meta=dict(workload="p9_single", dt=0.5))


def sweep_tp_counts(S=4096):
    rng = _rng(3)
    n = np.rint(rng.lognormal(np.log(SWEEP_MEDIAN), SWEEP_SIGMA, S))
    return np.clip(n, *SWEEP_CLIP).astype(np.int64)


def sweep_p9_grid(S=4096):
    """Planet Nine of each system: a 16x16x4x4 grid over mass 5..20 Earth
    masses, a 300..900 AU, e 0.1..0.6, i 0..30 deg (cell centres), shuffled
    with the master seed; remaining angles uniform."""
    gm = np.linspace(5, 20, 16) * EARTH_MASS
    ga = np.linspace(300, 900, 16)
    ge = np.linspace(0.1, 0.6, 4)
    gi = np.radians(np.linspace(0, 30, 4))
    M, A, E, I = np.meshgrid(gm, ga, ge, gi, indexing="ij")
    grid = np.stack([M.ravel(), A.ravel(), E.ravel(), I.ravel()], 1)
    perm = _rng(4).permutation(len(grid))
    grid = grid[perm]
    reps = int(np.ceil(S / len(grid)))
    return np.tile(grid, (reps, 1))[:S]


def p9_sweep(S=4096, systems=None, tp_scale=1.0) -> mint.State:
    """The 4096-system P9 sweep (or the listed subset of its systems).
    tp_scale < 1 shrinks every system's TP count (tests only)."""
    counts_all = sweep_tp_counts(max(S, 1))
    grid = sweep_p9_grid(max(S, 1))
    systems = np.arange(S) if systems is None else np.asarray(systems)
    ms, xs, vs, txs, tvs, counts = [], [], [], [], [], []
    for s in systems:
        rng = _rng(5, int(s))
        gm, gel = _giants(rng)
        pm, pa, pe, pi = grid[s]
        ang = rng.uniform(0, 2 * np.pi, 3)
        m = np.concatenate([[pm], gm])
        els = [np.concatenate([[p], g]) for p, g in zip((pi, ang[0], ang[1], ang[2], pe, pa), gel)]
        m, x, v = _massive_state(m, els)
        n = max(1, int(round(counts_all[s] * tp_scale)))
        tx, tv = _tps(rng, n)
        ms.append(m)
        xs.append(x)
        vs.append(v)
        txs.append(tx)
        tvs.append(tv)
        counts.append(n)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return mint.State(off, np.array(ms), np.array(xs), np.array(vs), np.concatenate(txs), np.concatenate(tvs),
                      meta=dict(workload="p9_sweep", dt=0.5, systems=len(systems)))


def inner_belt(n_tp=20000, a_range=(1.5, 3.0), seed_key=0) -> mint.State:
    """Sweep system 0's massive bodies with TPs on inner orbits (a in a_range
    AU, e < 0.2, i < 10 deg).  At dt = 0.5 yr their local step tau =
    dt sqrt(mu/r^3) is 0.08-0.3, the regime where the Kepler solve needs more
    than one Newton iteration from the series guess."""
    base = p9_sweep(systems=[0], tp_scale=1 / 256)
    rng = _rng(6, seed_key)
    a = rng.uniform(*a_range, n_tp)
    e = rng.uniform(0.0, 0.2, n_tp)
    inc = np.radians(rng.uniform(0.0, 10.0, n_tp))
    ang = rng.uniform(0.0, 2 * np.pi, (3, n_tp))
    tx, tv = mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], e, a)
    return mint.State(np.array([0, n_tp], np.int64), base.m, base.mx, base.mv, tx, tv,
                      meta=dict(workload="inner_belt", dt=0.5))


def embryo_radii(m, density=EMBRYO_DENSITY):
    """Physical radius [AU] of a body of mass m [Msun] at a density [g/cm^3]."""
    rho = density * AU_CM ** 3 / MSUN_G          # Msun / AU^3
    return np.cbrt(3.0 * np.asarray(m, float) / (4.0 * np.pi * rho))


def embryos(n=1000, seed_key=0) -> mint.State:
    """Sun + n planetary embryos (encounter-heavy, all massive): m = 3e-7
    U[0.8, 1.2], a ~ U[0.5, 4] AU, e ~ Rayleigh(0.01), i ~ Rayleigh(0.005 rad),
    remaining angles uniform.  Physical radii from a density of 3 g/cm^3 go in
    meta["radius"]; the configuration integrates with dt 0.002 yr, r_c = 3 Hill
    radii and BS tolerance 1e-11 (meta)."""
    rng = _rng(7, seed_key)
    m = 3.0e-7 * rng.uniform(0.8, 1.2, n)
    a = rng.uniform(0.5, 4.0, n)
    e = rng.rayleigh(0.01, n)
    inc = rng.rayleigh(0.005, n)
    ang = rng.uniform(0.0, 2 * np.pi, (3, n))
    x, v = mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], e, a)
    return mint.State(np.array([0, 0], np.int64), m[None], x[None], v[None], np.zeros((0, 3)), np.zeros((0, 3)),
                      meta=dict(workload="embryos", dt=0.002, n_hill=3.0, bs_tol=1e-11,
                                radius=embryo_radii(m)[None]))
