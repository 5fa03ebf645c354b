"""Host-side checks that need no GPU: the C ABI library loads and exports
every entry point include/mint.h declares, the host helpers agree with the
oracle, the workload generators are deterministic, and the CLI rejects bad
input before touching a device."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("mint.h", "mint_flops.h")]
CLI = os.path.join(ROOT, "p9-integrator_amd", "mint")


def declared_symbols():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mint_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = mint.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert L.mint_abi_version() == 2


def test_python_mirror_binds_every_declared_symbol():
    assert set(declared_symbols()) == set(mint._SIGS)


def test_flop_counts_of_the_roofline():
    # the per-unit figures the bench prices the kernels with: 529 FLOP per
    # TP-step at 5 bodies and 3 Newton iterations, 28 per ordered massive pair
    assert mint.flops_per_tp_step(3.0, 5) == 529.0
    assert mint.flops_per_tp_step(1.0, 5) == 449.0
    assert mint.flops_per_tp_step(0.0, 2) == 112 + 24 + 18 + 60
    assert mint.flops_per_kick_pass(1000) == 28.0 * 1000 * 999


def test_no_device_is_a_clean_error():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    with pytest.raises(mint.MintError, match="no HIP device"):
        mint.Mint()


def test_hill_radii_match_oracle_bitwise():
    st = W.p9_sweep(systems=np.arange(8), tp_scale=1 / 64)
    a = mint.hill_radii(1.0, 3.0, st.m, st.mx, st.mv)
    b = O.hill_radii(1.0, 3.0, st.m, st.mx, st.mv)
    assert np.array_equal(a, b)


def test_elements_round_trip():
    rng = np.random.default_rng(1)
    n = 500
    inc = rng.uniform(0, 3.0, n)
    node = rng.uniform(0, 2 * np.pi, n)
    argp = rng.uniform(0, 2 * np.pi, n)
    M = rng.uniform(0, 2 * np.pi, n)
    e = rng.uniform(0, 0.95, n)
    a = rng.uniform(1, 1000, n)
    x, v = mint.elements_to_cartesian(1.0, inc, node, argp, M, e, a)
    i2, n2, w2, M2, e2, a2 = mint.cartesian_to_elements(1.0, x, v)
    assert np.allclose(a2, a, rtol=1e-11)
    assert np.allclose(e2, e, rtol=0, atol=1e-11)
    assert np.allclose(i2, inc, atol=1e-11)
    x3, v3 = mint.elements_to_cartesian(1.0, i2, n2, w2, M2, e2, a2)
    assert np.allclose(x3, x, rtol=0, atol=1e-9 * np.abs(x).max())
    # Kepler energy from elements
    E = 0.5 * np.sum(v * v, 1) - 1 / np.linalg.norm(x, axis=1)
    assert np.allclose(E, -0.5 / a, rtol=1e-12)


def test_hyperbolic_elements():
    x, v = mint.elements_to_cartesian(1.0, [0.3], [0.1], [0.2], [0.7], [1.4], [-12.0])
    E = 0.5 * v[0] @ v[0] - 1 / np.linalg.norm(x[0])
    assert E == pytest.approx(0.5 / 12.0, rel=1e-12)


def test_sweep_generator_is_seeded_and_sized():
    counts = W.sweep_tp_counts(4096)
    assert counts.min() >= 256 and counts.max() <= 8192
    assert 5.4e6 < counts.sum() < 5.8e6
    assert np.median(counts) == pytest.approx(1024, rel=0.05)
    a = W.p9_sweep(systems=[5, 9], tp_scale=1 / 32)
    b = W.p9_sweep(systems=[9], tp_scale=1 / 32)
    assert np.array_equal(a.mx[1], b.mx[0])
    assert np.array_equal(a.tx[a.tp_offset[1]:], b.tx)
    grid = W.sweep_p9_grid(4096)
    assert len(np.unique(grid, axis=0)) == 4096
    assert grid[:, 0].min() == pytest.approx(5 * W.EARTH_MASS) and grid[:, 1].max() == pytest.approx(900)


def _cli(args, tmp_path, text=None):
    cfg = tmp_path / "c.json"
    if text is not None:
        cfg.write_text(text)
    return subprocess.run([CLI] + [a.replace("@CFG", str(cfg)) for a in args], capture_output=True, text=True,
                          cwd=tmp_path, timeout=60)


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_usage_and_bad_json(tmp_path):
    r = _cli(["-t", "5"], tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr
    r = _cli(["-c", "@CFG"], tmp_path, '{"bodies": [ {"name": "x", "mass": 1e-3, ')
    assert r.returncode == 1 and "JSON parse error" in r.stderr
    r = _cli(["-c", "@CFG"], tmp_path, '{"bodies": [{"name": "x", "mass": 1e-3}]}')
    assert r.returncode == 1 and "semi_major_axis" in r.stderr
    body = dict(name="b", mass=1e-6, semi_major_axis=1.0, eccentricity=0.0, inclination=0.0,
                longitude_of_ascending_node=0.0, argument_of_perihelion=0.0, mean_anomaly=0.0)
    tp = dict(body, name="tp", mass=0.0, semi_major_axis=2.0)
    r = _cli(["-c", "@CFG"], tmp_path, json.dumps({"bodies": [body] * 9 + [tp]}))
    assert r.returncode == 1 and "massive" in r.stderr  # test particles need at most 8 massive bodies
    r = _cli(["-c", "@CFG"], tmp_path, json.dumps({"bodies": [body] * 2049}))
    assert r.returncode == 1 and "massive" in r.stderr
    r = _cli(["-c", "@CFG", "-t"], tmp_path, json.dumps({"bodies": [body]}))
    assert r.returncode == 2


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_rejects_missing_checkpoint_argument(tmp_path):
    body = dict(name="b", mass=1e-6, semi_major_axis=1.0, eccentricity=0.0, inclination=0.0,
                longitude_of_ascending_node=0.0, argument_of_perihelion=0.0, mean_anomaly=0.0)
    r = _cli(["-c", "@CFG", "-r"], tmp_path, json.dumps({"bodies": [body]}))
    assert r.returncode == 2 and "usage" in r.stderr
