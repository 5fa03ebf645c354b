"""The oracle reproduces the committed golden fixtures bit for bit."""
import os

import numpy as np
import pytest

import oracle as O
from golden_io import all_fixtures, load


@pytest.mark.parametrize("path", all_fixtures(), ids=os.path.basename)
def test_oracle_reproduces_golden(path):
    d, st, ref = load(path)
    out, status, stats = O.run(st, d["steps"], d["dt"], guess=d["guess"], bs_tol=d["bs_tol"], radii=d["radii"])
    for k in ("mx", "mv", "tx", "tv", "rc"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.array_equal(status, ref["status"])
    assert {k: stats[k] for k in d["stats"]} == d["stats"]


def test_golden_set_covers_the_encounter_path():
    stats = [load(p)[0]["stats"] for p in all_fixtures()]
    assert any(s["bs_steps"] > 0 for s in stats)
    assert any(s["kepler_solves"] > 0 for s in stats)
    # massive encounter groups, including groups of three or more
    assert any(s.get("massive_groups", 0) > 0 for s in stats)
    assert any(s.get("massive_group_bodies", 0) > 2 * s.get("massive_groups", 0) > 0 for s in stats)
