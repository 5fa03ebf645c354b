"""Seeded configuration writer in the reference's JSON schema (the input of the
`mint` command line tool).

The reference generates its configurations with utils/sim_config.py, from an
unseeded RNG, writing fixed file names at import time.  This module builds the
same structures from a seed and returns them (or writes them where asked):

  p9_config(...)        Planet Nine set-up (utils/sim_config.py:69-207): 32
                        Planet Nine sweeps on a linear-shuffled combinatorial
                        grid, the fixed Planet Nine, the four giants with the
                        reference's elements, and the TNO ladder
  possible_orbits(...)  the 5000-entry candidate list written beside it
  embryo_config(...)    planetary embryos (utils/sim_config.py:32-63)
  earth_mars_config()   Sun + Earth-like + Mars-like (the two-planet case)
  embryos_1000_config() 1000 embryos with density 3 g/cm^3, dt 0.002, r_c = 3
                        Hill radii, BS tolerance 1e-11
  sweep_grid_config()   giants + Planet Nine sweeps over the 16x16x4x4 grid of
                        the 4096-system sweep (one shared TP set per file, as
                        the schema allows)

Angles are radians, masses solar masses, lengths AU.  Massive bodies (mass >
1e-12) must come first in "bodies" and a sweep replaces body 0, mass
included (src/mercurius.cu:4-15,126-164).

    python mint_configs.py p9 out.json [--seed S]
"""
from __future__ import annotations

import argparse
import json

import numpy as np

import mint_workloads as W

EARTH_MASS = 3.00338e-06            # utils/sim_config.py:91


def linear_shuffled(bounds, n, rng):
    """n values lower + i (upper - lower)/n, i = 0..n-1, in a seeded random
    order (generate_sweep_array's linear schedule, utils/sim_config.py:8-19)."""
    lower, upper = bounds
    spacing = (upper - lower) / n
    vals = [lower + i * spacing for i in range(n)]
    order = rng.permutation(n)
    return [vals[k] for k in order]


def combinatorial_sweep(n, bounds_list, rng):
    """n parameter vectors; parameter p of vector j is the j-th entry of its
    own shuffled linear ladder (combinatorial_sweep, utils/sim_config.py:21-29)."""
    cols = [linear_shuffled(b, n, rng) for b in bounds_list]
    return [[cols[p][j] for p in range(len(cols))] for j in range(n)]


def _range(mean, up, down):
    return [mean - down, mean + up]


def _body(name, m, a, e, i, node, argp, manom):
    return {"name": name, "mass": float(m), "semi_major_axis": float(a), "eccentricity": float(e),
            "inclination": float(i), "longitude_of_ascending_node": float(node),
            "argument_of_perihelion": float(argp), "mean_anomaly": float(manom)}


GIANTS = [  # name, mass, a, e, i [deg], node [deg], argp [deg] (utils/sim_config.py:150-193)
    ("Neptune", 5.14855965e-5, 30.06896348, 0.00858587, 1.76917, 131.72169, -86.75034),
    ("Uranus", 4.36430044e-5, 19.19126393, 0.04716771, 0.76986, 74.22988, -96.73436),
    ("Saturn", 0.000285716656, 9.53707032, 0.05415060, 2.48446, 113.71504, -21.2831),
    ("Jupiter", 0.000954588, 5.20336301, 0.04839266, 1.30530, 100.55615, -85.8023),
]


def p9_config(seed=0, n_tnos=640 * 16 - 5, n_sweeps=32):
    rng = np.random.default_rng([W.MASTER_SEED, 10, seed])
    a_lo, a_hi = 150.0, 500.0
    q_lo, q_hi = 30.0, 50.0
    e_lo, e_hi = 1.0 - q_lo / a_lo, 1.0 - q_hi / a_hi
    a_sp = (a_hi - a_lo) / n_tnos
    e_sp = (e_hi - e_lo) / n_tnos
    i_sp = np.radians(25.0) / n_tnos
    bounds = [_range(6.6 * EARTH_MASS, 2.6 * EARTH_MASS, 1.7 * EARTH_MASS), _range(500, 170, 120),
              _range(np.radians(15.6), np.radians(5.2), np.radians(5.4)), _range(0.26, 0.1, 0.1),
              _range(np.radians(246.7), np.radians(15.1), np.radians(13.4)),
              _range(np.radians(96.9), np.radians(17.3), np.radians(15.5))]
    sweeps = []
    for m, a, inc, e, varpi, node in combinatorial_sweep(n_sweeps, bounds, rng):
        # longitude of perihelion -> argument of perihelion
        sweeps.append(_body("Planet Nine", m, a, e, inc, node, varpi - node, 0.0))
    bodies = [_body("Planet Nine", 5 * EARTH_MASS, 500, 0.25, np.radians(20), np.radians(86), np.radians(138),
                    2 * np.pi + np.radians(-176.2))]
    giants = [_body(n, m, a, e, np.radians(i), np.radians(o), 2 * np.pi + np.radians(w), 0.0)
              for n, m, a, e, i, o, w in GIANTS]
    tnos = [_body(f"TNO-{k}", 0.0, a_lo + a_sp * k, e_lo + k * e_sp, i_sp * k, *rng.uniform(0, 2 * np.pi, 3))
            for k in range(n_tnos)]
    rest = sorted(giants + tnos, key=lambda b: b["mass"], reverse=True)
    return {"sweeps": sweeps, "bodies": bodies + rest}


def possible_orbits(seed=0, n=5000):
    rng = np.random.default_rng([W.MASTER_SEED, 11, seed])
    bounds = [_range(500, 170, 120), _range(np.radians(15.6), np.radians(5.2), np.radians(5.4)),
              _range(0.26, 0.1, 0.1), _range(np.radians(246.7), np.radians(15.1), np.radians(13.4)),
              _range(np.radians(96.9), np.radians(17.3), np.radians(15.5))]
    return [{"semi_major_axis": a, "inclination": i, "eccentricity": e, "argument_of_periapsis": w,
             "ascending_node": o} for a, i, e, w, o in combinatorial_sweep(n, bounds, rng)]


def embryo_config(seed=0, n=30):
    """Planetary embryos: masses 0.6 lunar .. 0.2 Earth masses, a 0.5..1.2 AU
    and e 0..0.01 on linear ladders, planar, random angles, heaviest first."""
    rng = np.random.default_rng([W.MASTER_SEED, 12, seed])
    m_hi = 0.2 / 332946
    m_lo = 0.6 / 27068510
    bodies = [_body(f"Small-Body-{k}", m_lo + (m_hi - m_lo) / n * k, 0.5 + (1.2 - 0.5) / n * k, k * 0.01 / n, 0.0,
                    *rng.uniform(0, 2 * np.pi, 3)) for k in range(n)]
    bodies.sort(key=lambda b: b["mass"], reverse=True)
    return {"bodies": bodies}


def _from_state(st, names=None, tp_names=None):
    """Reference-schema bodies of system 0 of a mint.State (heliocentric)."""
    import mint

    out = []
    el = mint.cartesian_to_elements(1.0, st.mx[0], st.mv[0])
    for k in range(st.n_massive):
        inc, node, argp, manom, e, a = (float(v[k]) for v in el)
        out.append(_body(names[k] if names else f"Body-{k}", st.m[0, k], a, e, inc, node, argp, manom))
    if st.n_tp:
        el = mint.cartesian_to_elements(1.0, st.tx, st.tv)
        for k in range(st.n_tp):
            inc, node, argp, manom, e, a = (float(v[k]) for v in el)
            out.append(_body(f"TP-{k}", 0.0, a, e, inc, node, argp, manom))
    return out


def earth_mars_config():
    st = W.earth_mars()
    return {"bodies": _from_state(st, ["Earth", "Mars"]), "dt": 1e-3}


def embryos_1000_config(seed=0):
    st = W.embryos(1000, seed_key=seed)
    return {"bodies": _from_state(st), "dt": 0.002, "n_hill": 3.0, "bs_tol": 1e-11,
            "density": W.EMBRYO_DENSITY}


def sweep_grid_config(n_sweeps=4096, n_tp=1024, seed=0):
    """Giants of sweep system 0, n_tp TPs (a 150..1000, q 30..50), and one
    Planet Nine sweep per grid point (mass included)."""
    st = W.p9_sweep(systems=[0], tp_scale=n_tp / W.sweep_tp_counts(4096)[0])
    base = _from_state(st, ["Planet Nine", "Jupiter", "Saturn", "Uranus", "Neptune"])
    grid = W.sweep_p9_grid(n_sweeps)
    rng = np.random.default_rng([W.MASTER_SEED, 13, seed])
    sweeps = [_body("Planet Nine", m, a, e, i, *rng.uniform(0, 2 * np.pi, 2), 0.0) for m, a, e, i in grid]
    return {"sweeps": sweeps, "bodies": base, "dt": 0.5}


BUILDERS = {"p9": p9_config, "embryos": embryo_config, "earth_mars": earth_mars_config,
            "embryos_1000": embryos_1000_config, "sweep_grid": sweep_grid_config}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("which", choices=sorted(BUILDERS) + ["possible_orbits"])
    ap.add_argument("out")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    if a.which == "possible_orbits":
        obj = possible_orbits(a.seed)
    elif a.which in ("earth_mars",):
        obj = BUILDERS[a.which]()
    else:
        obj = BUILDERS[a.which](seed=a.seed)
    with open(a.out, "w") as f:
        json.dump(obj, f, indent=1)


if __name__ == "__main__":
    main()
