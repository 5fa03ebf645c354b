"""Load the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

import mint

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _arr(h, shape=None):
    a = np.array([float.fromhex(x) for x in h], dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def load(path):
    with open(path) as f:
        d = json.load(f)
    i = d["input"]
    S, NM = i["S"], i["NM"]
    st = mint.State(np.array(i["tp_offset"], np.int64), _arr(i["m"], (S, NM)), _arr(i["mx"], (S, NM, 3)),
                    _arr(i["mv"], (S, NM, 3)), _arr(i["tx"], (-1, 3)), _arr(i["tv"], (-1, 3)))
    o = d["output"]
    out = dict(mx=_arr(o["mx"], (S, NM, 3)), mv=_arr(o["mv"], (S, NM, 3)), tx=_arr(o["tx"], (-1, 3)),
               tv=_arr(o["tv"], (-1, 3)), rc=_arr(o["rc"], (S, NM)), status=np.array(o["status"], np.uint32))
    d.setdefault("bs_tol", 1e-12)
    d["radii"] = None if d.get("radii") is None else _arr(d["radii"], (S, NM))
    return d, st, out


def all_fixtures():
    return sorted(glob.glob(os.path.join(GOLDEN, "*.json")))
