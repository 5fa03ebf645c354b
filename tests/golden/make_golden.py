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
from gpu_cases import encounter_system  # noqa: E402


def cases():
    yield "earth_mars_100", W.earth_mars(), 100, 1e-3
    p = W.p9_single(64)
    yield "p9_single_64tp_100", p, 100, 0.5
    yield "p9_sweep_2sys_100", W.p9_sweep(systems=[0, 1], tp_scale=1 / 16), 100, 0.5
    yield "encounter_cloud_40", encounter_system(n_tp=96, seed=2), 40, 0.5


def hexs(a):
    return [float(x).hex() for x in np.asarray(a).ravel()]


def state_dict(st):
    return dict(tp_offset=st.tp_offset.tolist(), m=hexs(st.m), mx=hexs(st.mx), mv=hexs(st.mv), tx=hexs(st.tx),
                tv=hexs(st.tv), S=int(st.m.shape[0]), NM=int(st.m.shape[1]))


def main():
    for name, st, n, dt in cases():
        out, status, stats = O.run(st, n, dt, guess=1)
        d = dict(name=name, steps=n, dt=dt, guess=1, input=state_dict(st),
                 output=dict(mx=hexs(out["mx"]), mv=hexs(out["mv"]), tx=hexs(out["tx"]), tv=hexs(out["tv"]),
                             rc=hexs(out["rc"]), status=status.tolist()),
                 stats=stats)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(d, f)
        print(name, stats)


if __name__ == "__main__":
    main()
