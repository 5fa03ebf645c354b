"""Shared GPU-test inputs: sweep slices and an encounter-heavy system."""
import numpy as np

import mint
import mint_workloads as W


def encounter_system(n_tp=512, seed=11, sys_index=0):
    """Sweep system `sys_index` with its TPs replaced by a cloud around Planet
    Nine (body 0) and Neptune (body 4): offsets 0.05-1.5 r_c, velocities the
    body's own plus a small random part, so most TPs are inside a changeover
    zone for many steps."""
    st = W.p9_sweep(systems=[sys_index], tp_scale=1 / 64)
    rc = mint.hill_radii(1.0, 3.0, st.m[0], st.mx[0], st.mv[0])
    rng = np.random.default_rng(seed)
    tx, tv = [], []
    for i in range(n_tp):
        j = 0 if i % 2 == 0 else 4
        d = rng.normal(size=3)
        d *= rc[j] * rng.uniform(0.05, 1.5) / np.linalg.norm(d)
        dv = rng.normal(size=3) * 0.02 * np.linalg.norm(st.mv[0, j])
        tx.append(st.mx[0, j] + d)
        tv.append(st.mv[0, j] + dv)
    st.tx = np.array(tx)
    st.tv = np.array(tv)
    st.tp_offset = np.array([0, n_tp], np.int64)
    return st


def planet_cloud(nm, n_tp=96, seed=0):
    """nm planets (masses 1e-4 (1 + j) Msun, a = 5 x 1.6^j AU, e 0.03, small
    inclinations) and a cloud of TPs around the first and the last planet
    (offsets 0.05-1.5 r_c, velocities the planet's own plus a small random
    part): close encounters at every massive count 1..8."""
    rng = np.random.default_rng([20261203, 10, nm, seed])
    m = 1e-4 * (1.0 + np.arange(nm))
    a = 5.0 * 1.6 ** np.arange(nm)
    inc = rng.uniform(0.0, 0.02, nm)
    ang = rng.uniform(0.0, 2 * np.pi, (3, nm))
    x, v = mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], np.full(nm, 0.03), a)
    rc = mint.hill_radii(1.0, 3.0, m, x, v)
    tx, tv = [], []
    for i in range(n_tp):
        j = 0 if i % 2 == 0 else nm - 1
        d = rng.normal(size=3)
        d *= rc[j] * rng.uniform(0.05, 1.5) / np.linalg.norm(d)
        tx.append(x[j] + d)
        tv.append(v[j] + rng.normal(size=3) * 0.02 * np.linalg.norm(v[j]))
    return mint.State(np.array([0, n_tp], np.int64), m[None], x[None], v[None], np.array(tx), np.array(tv),
                      meta=dict(workload="planet_cloud", nm=nm))


def unbound_cloud(n=256, seed=17):
    """The giants of sweep system 0 and n TPs at 20-200 AU in four sets of
    speed, as multiples of the escape speed: hyperbolic (1.5-20), near the
    parabolic limit (1 +- 1e-9 .. 1e-3), around escape (0.9-1.1) and slow
    (0.05-0.3); every fourth TP moves nearly radially."""
    st = W.p9_sweep(systems=[0], tp_scale=1 / 64).copy()
    rng = np.random.default_rng([20261203, 10, seed])
    q = n // 4
    r = rng.uniform(20.0, 200.0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    radial = np.arange(n) % 4 == 3
    w[radial] = u[radial] + 1e-3 * w[radial]
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    sign = np.where(rng.uniform(size=q) < 0.5, -1.0, 1.0)
    f = np.concatenate([rng.uniform(1.5, 20.0, q), 1.0 + sign * 10.0 ** rng.uniform(-9, -3, q),
                        rng.uniform(0.9, 1.1, q), rng.uniform(0.05, 0.3, n - 3 * q)])
    v = f * np.sqrt(2.0 / r)  # central mass 1, G = 1
    st.tx = u * r[:, None]
    st.tv = w * v[:, None]
    st.tp_offset = np.array([0, n], np.int64)
    return st


def run_gpu(st, nsteps, dt, form, guess, stats=1, **kw):
    with mint.Mint(dt=dt, form=form, kepler_guess=guess, collect_stats=stats, **kw) as g:
        g.upload(st)
        g.step(nsteps)
        return g.download(), g.status(), g.stats()


def embryo_ring(n=64, seed=0, a_lo=1.0, a_hi=1.02, mass=3.0e-7):
    """n embryos packed into a narrow ring (a in [a_lo, a_hi] AU, e and i
    Rayleigh 0.01 / 0.005): dense enough that several encounter groups, some
    of three or more bodies, form within tens of steps at dt = 0.002."""
    rng = np.random.default_rng([20261203, 8, seed])
    m = mass * rng.uniform(0.8, 1.2, n)
    a = rng.uniform(a_lo, a_hi, n)
    e = rng.rayleigh(0.01, n)
    inc = rng.rayleigh(0.005, n)
    ang = rng.uniform(0.0, 2 * np.pi, (3, n))
    x, v = mint.elements_to_cartesian(1.0, inc, ang[0], ang[1], ang[2], e, a)
    return mint.State(np.array([0, 0], np.int64), m[None], x[None], v[None], np.zeros((0, 3)), np.zeros((0, 3)),
                      meta=dict(workload="embryo_ring", dt=0.002, radius=W.embryo_radii(m)[None]))


def embryo_cluster(n=8, seed=0, span=0.06, a_lo=1.0, a_hi=1.004, mass=3.0e-7):
    """n embryos within `span` radians of orbital longitude and a in
    [a_lo, a_hi]: they start in and out of each other's changeover zones, so
    groups form, merge and split over a few hundred steps."""
    rng = np.random.default_rng([20261203, 9, seed])
    m = mass * rng.uniform(0.8, 1.2, n)
    a = rng.uniform(a_lo, a_hi, n)
    e = rng.rayleigh(0.002, n)
    inc = rng.rayleigh(0.001, n)
    z = np.zeros(n)
    x, v = mint.elements_to_cartesian(1.0, inc, z, z, rng.uniform(0.0, span, n), e, a)
    return mint.State(np.array([0, 0], np.int64), m[None], x[None], v[None], np.zeros((0, 3)), np.zeros((0, 3)),
                      meta=dict(workload="embryo_cluster", dt=0.002, radius=W.embryo_radii(m)[None]))


def coorbital_system(n_tp=64, seed=5):
    """Sweep system 0 with Neptune (body 4) replaced by a pair of 1e-5 Msun
    co-orbitals 0.02 rad apart at its distance, which start inside their
    changeover radius and stay near it (a massive encounter group beside three
    free bodies), plus TPs around the pair."""
    st = W.p9_sweep(systems=[0], tp_scale=1 / 64)
    m = np.concatenate([st.m[0, :4], [1e-5, 1.2e-5]])
    a = 30.1
    phase = np.random.default_rng(seed).uniform(0, 2 * np.pi)
    els = [np.array([0.001, 0.002]), np.array([0.3, 0.3]), np.array([0.1, 0.1]),
           np.array([phase, phase + 0.02]), np.array([0.001, 0.002]), np.array([a, a])]
    x2, v2 = mint.elements_to_cartesian(1.0, *els)
    mx = np.concatenate([st.mx[0, :4], x2])
    mv = np.concatenate([st.mv[0, :4], v2])
    rc = mint.hill_radii(1.0, 3.0, m, mx, mv)
    rng = np.random.default_rng(seed + 1)
    tx, tv = [], []
    for i in range(n_tp):
        j = 4 + (i % 2)
        d = rng.normal(size=3)
        d *= rc[j] * rng.uniform(0.2, 2.0) / np.linalg.norm(d)
        tx.append(mx[j] + d)
        tv.append(mv[j] + rng.normal(size=3) * 0.01 * np.linalg.norm(mv[j]))
    return mint.State(np.array([0, n_tp], np.int64), m[None], mx[None], mv[None], np.array(tx), np.array(tv),
                      meta=dict(workload="coorbital", dt=0.5))
