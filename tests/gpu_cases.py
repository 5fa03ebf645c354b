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


def run_gpu(st, nsteps, dt, form, guess, stats=1, **kw):
    with mint.Mint(dt=dt, form=form, kepler_guess=guess, collect_stats=stats, **kw) as g:
        g.upload(st)
        g.step(nsteps)
        return g.download(), g.status(), g.stats()
