"""Seeded configuration writer (reference JSON schema) and the CLI's reading
of it: structure of the Planet Nine set-up (linear-shuffled sweep grid, body
order), the embryo set, the C-configurations, and a dry run of the `mint`
tool on generated configurations and, where the reference checkout is
present, on the reference's own example files."""
import json
import os
import subprocess

import numpy as np
import pytest

import mint_configs as MC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "p9-integrator_amd", "mint")
REF_EXAMPLES = "/root/reference/examples"
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


def _ladder(lo, hi, n):
    sp = (hi - lo) / n
    return [lo + i * sp for i in range(n)]


def test_linear_shuffled_is_a_permutation_of_the_ladder():
    rng = np.random.default_rng(0)
    v = MC.linear_shuffled((2.0, 5.0), 37, rng)
    assert sorted(v) == _ladder(2.0, 5.0, 37)
    assert v != _ladder(2.0, 5.0, 37)


def test_p9_config_structure():
    c = MC.p9_config(seed=1)
    assert len(c["sweeps"]) == 32
    names = [b["name"] for b in c["bodies"][:5]]
    assert names == ["Planet Nine", "Jupiter", "Saturn", "Neptune", "Uranus"]
    tnos = c["bodies"][5:]
    assert len(tnos) == 640 * 16 - 5 and all(b["mass"] == 0 for b in tnos)
    a = [b["semi_major_axis"] for b in tnos]
    assert a[0] == 150.0 and np.all(np.diff(a) > 0) and a[-1] < 500.0
    # a and e run on linear ladders between (150 AU, q 30) and (500 AU, q 50),
    # so q peaks at 51.6 AU in between (as in the reference's set-up)
    q = [b["semi_major_axis"] * (1 - b["eccentricity"]) for b in tnos]
    assert 29.99 < min(q) and max(q) < 51.7
    E = 3.00338e-06
    assert sorted(s["mass"] for s in c["sweeps"]) == _ladder(6.6 * E - 1.7 * E, 6.6 * E + 2.6 * E, 32)
    assert sorted(s["semi_major_axis"] for s in c["sweeps"]) == _ladder(380, 670, 32)
    varpi = sorted(s["argument_of_perihelion"] + s["longitude_of_ascending_node"] for s in c["sweeps"])
    assert np.allclose(varpi, _ladder(np.radians(246.7 - 13.4), np.radians(246.7 + 15.1), 32), rtol=0, atol=1e-12)
    assert all(s["mean_anomaly"] == 0 for s in c["sweeps"])


def test_configs_are_seeded():
    assert MC.p9_config(seed=3, n_tnos=50) == MC.p9_config(seed=3, n_tnos=50)
    a, b = MC.p9_config(seed=3, n_tnos=50), MC.p9_config(seed=4, n_tnos=50)
    assert a != b and sorted(s["mass"] for s in a["sweeps"]) == sorted(s["mass"] for s in b["sweeps"])
    assert MC.embryo_config(seed=2) == MC.embryo_config(seed=2)
    assert len(MC.possible_orbits(seed=1)) == 5000


def test_embryo_config_heaviest_first():
    c = MC.embryo_config()
    m = [b["mass"] for b in c["bodies"]]
    assert len(m) == 30 and m == sorted(m, reverse=True)
    assert m[0] < 0.2 / 332946 and m[-1] == pytest.approx(0.6 / 27068510, rel=1e-15)


def _dry(path, *flags):
    r = subprocess.run([CLI, "-c", path, "-n", *flags], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.split(": ", 1)[1]


@needs_cli
def test_cli_reads_generated_configs(tmp_path):
    cases = {"p9": ("32 system(s), 5 massive and 10235 massless", MC.p9_config()),
             "embryos": ("1 system(s), 30 massive and 0 massless", MC.embryo_config()),
             "embryos_1000": ("1 system(s), 1000 massive and 0 massless", MC.embryos_1000_config()),
             "earth_mars": ("1 system(s), 2 massive and 0 massless", MC.earth_mars_config()),
             "sweep_grid": ("4096 system(s), 5 massive and 1024 massless", MC.sweep_grid_config())}
    for name, (expect, cfg) in cases.items():
        p = tmp_path / f"{name}.json"
        p.write_text(json.dumps(cfg))
        assert _dry(str(p), "-a").startswith(expect), name
        # the reference's per-device slice: sweeps [4 d, 4 d + 4)
        n = int(expect.split()[0])
        assert _dry(str(p)).startswith(f"{min(n, 4)} system(s)"), name
    last = (32 - 1) // 4
    assert _dry(str(tmp_path / "p9.json"), "-d", str(last)).startswith("4 system(s)")
    r = subprocess.run([CLI, "-c", str(tmp_path / "p9.json"), "-n", "-d", str(last + 1)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "selects sweeps 32..35" in r.stderr
    assert "dt 0.002 yr, n_hill 3, bs_tol 1e-11" in _dry(str(tmp_path / "embryos_1000.json"))


@needs_cli
@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference checkout not present")
def test_cli_reads_the_reference_examples_like_the_generated_ones(tmp_path):
    p = tmp_path / "p9.json"
    p.write_text(json.dumps(MC.p9_config()))
    # same shape and the same total massive mass (the sweep mass ladder does not
    # depend on the shuffle)
    assert _dry(os.path.join(REF_EXAMPLES, "p9.json"), "-a") == _dry(str(p), "-a")
    assert _dry(os.path.join(REF_EXAMPLES, "p9.json")).startswith("4 system(s), 5 massive")
    e = tmp_path / "e.json"
    e.write_text(json.dumps(MC.embryo_config()))
    ref_e = _dry(os.path.join(REF_EXAMPLES, "planetary_embryos.json"))
    assert ref_e.startswith("1 system(s), 30 massive") and _dry(str(e)).split(", total")[0] == ref_e.split(", total")[0]
    assert _dry(os.path.join(REF_EXAMPLES, "earth_and_mars.json")).startswith("1 system(s), 2 massive")
