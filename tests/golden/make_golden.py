"""Writes the golden fixtures: final FP64 states of the oracle on small seeded
workloads, stored bit-exactly (hex floats).  The reference publishes no
trajectories and cannot be built, so these pin the oracle against regressions
and give the GPU tests a fixed target; the oracle itself is pinned by
tests/test_oracle.py.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "p9-integrator_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(HERE)]

import mint_workloads as W  # noqa: E402
import oracle as O  # noqa: E402
from gpu_cases import coorbital_system, embryo_cluster, embryo_ring, encounter_system  # noqa: E402


def cases():
    """(name, state, steps, dt, BS tolerance, physical radii or None)"""
    yield "earth_mars_100", W.earth_mars(), 100, 1e-3, 1e-12, None
    p = W.p9_single(64)
    yield "p9_single_64tp_100", p, 100, 0.5, 1e-12, None
    yield "p9_sweep_2sys_100", W.p9_sweep(systems=[0, 1], tp_scale=1 / 16), 100, 0.5, 1e-12, None
    yield "encounter_cloud_40", encounter_system(n_tp=96, seed=2), 40, 0.5, 1e-12, None
    # massive encounter groups: beside TPs, in a dense 8-body cluster, and in a
    # 64-embryo ring (the many-body engine) with physical radii
    yield "coorbital_40", coorbital_system(), 40, 0.5, 1e-12, None
    c = embryo_cluster(8, seed=0)
    yield "embryo_cluster_8_400", c, 400, 0.002, 1e-11, None
    r = embryo_ring(64)
    yield "embryo_ring_64_300", r, 300, 0.002, 1e-11, r.meta["radius"]


def hexs(a):
    return [float(x).hex() for x in np.asarray(a).ravel()]


def state_dict(st):
    return dict(tp_offset=st.tp_offset.tolist(), m=hexs(st.m), mx=hexs(st.mx), mv=hexs(st.mv), tx=hexs(st.tx),
                tv=hexs(st.tv), S=int(st.m.shape[0]), NM=int(st.m.shape[1]))


def main():
    for name, st, n, dt, tol, rad in cases():
        out, status, stats = O.run(st, n, dt, guess=1, bs_tol=tol, radii=rad)
        d = dict(name=name, steps=n, dt=dt, guess=1, bs_tol=tol, radii=None if rad is None else hexs(rad),
                 input=state_dict(st),
                 output=dict(mx=hexs(out["mx"]), mv=hexs(out["mv"]), tx=hexs(out["tx"]), tv=hexs(out["tv"]),
                             rc=hexs(out["rc"]), status=status.tolist()),
                 stats=stats)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(d, f)
        print(name, stats)


if __name__ == "__main__":
    main()
