"""Full-size parity.  TPs are massless, so a subset of a system's TPs run by
the oracle must match, bit for bit, the same TPs inside the GPU run of the
whole system.  This checks the configurations at their real sizes:
  earth_mars   1e5 steps, dt 1e-3 (energy tracks the oracle exactly)
  p9_single    1e5 TPs, 1e4 steps, dt 0.5
  p9_sweep     4096 systems (5.49 M TPs), 2000 steps, learned+batched form
plus size-independent properties of the full runs (finite states, status)."""
import json
import os
import subprocess

import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
import train_learned as T

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _subset_tps(st, system, idx):
    sub = st.subset([system])
    sub.tx = sub.tx[idx]
    sub.tv = sub.tv[idx]
    sub.tp_offset = np.array([0, len(idx)], np.int64)
    return sub


def test_earth_mars_1e5_steps_energy_tracks_oracle():
    st = W.earth_mars()
    print("earth_mars 1e5 steps", flush=True)
    with mint.Mint(dt=1e-3, form=mint.FORM_BATCHED, chunk_steps=4096) as g:
        g.upload(st)
        E0 = g.energy()
        g.step(100_000)
        out = g.download()
        E1 = g.energy()
    ref, _, _ = O.run(st, 100_000, 1e-3, guess=1)
    assert np.array_equal(out.mx, ref["mx"]) and np.array_equal(out.mv, ref["mv"])
    dE = abs((E1[0] - E0[0]) / E0[0])
    assert dE < 1e-9, dE
    Eo = O.energy(st.m, ref["mx"], ref["mv"])
    assert abs(E1[0] - Eo[0]) <= 1e-15 * abs(E0[0])


def test_p9_single_full_size_subset_bitwise():
    st = W.p9_single(100_000)
    print("p9_single 1e5 TPs x 1e4 steps", flush=True)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, collect_stats=1) as g:
        g.upload(st)
        g.step(10_000)
        out = g.download()
        status = g.status()
        s = g.stats()
    assert np.all(np.isfinite(out.tx)) and np.all(np.isfinite(out.tv))
    assert not (status[0] & (mint.ST_NAN | mint.ST_NEWTON))
    idx = np.random.default_rng(1).choice(st.n_tp, 192, replace=False)
    ref, _, rs = O.run(_subset_tps(st, 0, idx), 10_000, 0.5, guess=1)
    assert np.array_equal(out.mx, ref["mx"])
    assert np.array_equal(out.tx[idx], ref["tx"]) and np.array_equal(out.tv[idx], ref["tv"])
    print(f"p9_single stats {s}", flush=True)


def test_p9_sweep_full_size_learned_batched_subset_bitwise():
    st = W.p9_sweep()
    print(f"p9_sweep full: {st.n_tp} TPs, 2000 steps", flush=True)
    with mint.Mint(dt=0.5, form=mint.FORM_BATCHED, kepler_guess=mint.GUESS_LEARNED, use_scorer=1,
                   chunk_steps=T.window(), collect_stats=1) as g:
        T.apply(g)
        g.upload(st)
        g.step(2000)
        out = g.download()
        status = g.status()
        s = g.stats()
    assert np.all(np.isfinite(out.tx))
    assert not np.any(status & (mint.ST_NAN | mint.ST_NEWTON | mint.ST_BS_DEPTH))
    assert s["bs_steps"] > 0 and s["prebucketed"] > 0
    rng = np.random.default_rng(2)
    for sysid in rng.choice(4096, 6, replace=False):
        a, b = st.tp_offset[sysid], st.tp_offset[sysid + 1]
        idx = rng.choice(b - a, min(24, b - a), replace=False)
        ref, _, _ = O.run(_subset_tps(st, sysid, idx), 2000, 0.5, guess=1)
        assert np.array_equal(out.mx[sysid], ref["mx"][0])
        assert np.array_equal(out.tx[a + idx], ref["tx"]) and np.array_equal(out.tv[a + idx], ref["tv"])


def test_cli_runs_a_reference_schema_config(tmp_path):
    cli = os.path.join(ROOT, "p9-integrator_amd", "mint")
    bodies = [dict(name="Jupiter", mass=9.54e-4, semi_major_axis=5.2, eccentricity=0.048, inclination=0.02,
                   longitude_of_ascending_node=1.75, argument_of_perihelion=4.78, mean_anomaly=0.3),
              dict(name="Saturn", mass=2.86e-4, semi_major_axis=9.55, eccentricity=0.054, inclination=0.04,
                   longitude_of_ascending_node=1.98, argument_of_perihelion=5.92, mean_anomaly=1.1)]
    bodies += [dict(name=f"TNO-{i}", mass=0, semi_major_axis=40 + i, eccentricity=0.1, inclination=0.1,
                    longitude_of_ascending_node=0.1 * i, argument_of_perihelion=0.2, mean_anomaly=0.3 * i)
               for i in range(5)]
    sweeps = [dict(name="Planet Nine", mass=1.5e-5 * (1 + k), semi_major_axis=500 + 10 * k, eccentricity=0.25,
                   inclination=0.3, longitude_of_ascending_node=1.5, argument_of_perihelion=2.6, mean_anomaly=0)
              for k in range(3)]
    cfg = dict(bodies=[dict(sweeps[0], name="Planet Nine")] + bodies, sweeps=sweeps, dt=0.5)
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    r = subprocess.run([cli, "-c", "c.json", "-t", "40", "-d", "0", "-i"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Energy drift" in r.stdout and "NaN counter after step 40: 0" in r.stdout
    text = (tmp_path / "output_0").read_text()
    assert text.startswith("#\n$\n") and text.count("$\n") == 3
    lines = [ln for ln in text.split("\n") if ln and ln not in ("#", "$")]
    assert len(lines) == 3 * 8 and all(len(ln.split()) == 3 for ln in lines)
    el = (tmp_path / "elements_0.txt").read_text()
    assert el.startswith("Timestep: 40") and el.count("Experiment:") == 3


def test_cli_embryos_with_density_and_resume(tmp_path):
    """An all-massive configuration through the CLI (40 embryos, radii from a
    density): a 40-step run equals a 20-step run saved with -s and resumed
    with -r for 20 more, to the printed digit, and the step count continues."""
    cli = os.path.join(ROOT, "p9-integrator_amd", "mint")
    rng = np.random.default_rng(7)
    bodies = [dict(name=f"E{i}", mass=3e-7 * rng.uniform(0.8, 1.2), semi_major_axis=rng.uniform(1.0, 1.05),
                   eccentricity=rng.rayleigh(0.01), inclination=rng.rayleigh(0.005),
                   longitude_of_ascending_node=rng.uniform(0, 6.28), argument_of_perihelion=rng.uniform(0, 6.28),
                   mean_anomaly=rng.uniform(0, 6.28)) for i in range(40)]
    (tmp_path / "e.json").write_text(json.dumps(dict(bodies=bodies, dt=0.002, bs_tol=1e-11, density=3.0)))

    def run(*a):
        r = subprocess.run([cli, "-c", "e.json", *a], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        return r.stdout

    run("-t", "40", "-o", "full")
    run("-t", "20", "-o", "part1", "-s", "ck.bin")
    out = run("-t", "20", "-o", "part2", "-r", "ck.bin", "-i")
    assert "Resumed from ck.bin at step 20" in out and "NaN counter after step 40: 0" in out
    assert (tmp_path / "full").read_text() == (tmp_path / "part2").read_text()
    assert (tmp_path / "elements_0.txt").read_text().startswith("Timestep: 40")


def test_checkpoint_resume_is_bitwise(tmp_path):
    st = W.p9_sweep(systems=np.arange(5), tp_scale=1 / 4)
    path = str(tmp_path / "run.ckpt")
    kw = dict(dt=0.5, form=mint.FORM_BATCHED, chunk_steps=32, block_size=128, tps_per_lane=1)
    with mint.Mint(**kw) as g:
        g.upload(st)
        g.step(100)
        ref = g.download()
    with mint.Mint(**kw) as g:
        g.upload(st)
        g.step(60)
        g.save_checkpoint(path, 60)
    with mint.Mint(**kw) as g:
        k = g.load_checkpoint(path)
        assert k == 60
        g.step(40)
        out = g.download()
    assert np.array_equal(out.tx, ref.tx) and np.array_equal(out.tv, ref.tv)
    assert np.array_equal(out.mx, ref.mx) and np.array_equal(out.mv, ref.mv)


def test_embryos_1000_full_run_1e4_steps():
    """The all-massive embryo configuration at full length: 1000 bodies, 1e4
    steps of 0.002 yr on the many-body engine.  Size-independent checks: the
    state stays finite, no capacity limit or BS budget is hit, encounter
    groups formed, and the energy error stays bounded."""
    st = W.embryos(1000)
    with mint.Mint(dt=0.002, bs_tol=1e-11, collect_stats=1) as g:
        g.upload(st)
        g.set_radii(st.meta["radius"])
        E0 = g.energy()
        g.step(10000)
        E1 = g.energy()
        status = g.status()
        s = g.stats()
        out = g.download()
    print("embryos 1e4 steps:", s, "dE/E", (E1 - E0) / E0, flush=True)
    assert np.all(np.isfinite(out.mx)) and np.all(np.isfinite(out.mv))
    assert not status[0] & (mint.ST_NAN | mint.ST_NEWTON | mint.ST_BS_DEPTH | mint.ST_LIMIT)
    assert s["massive_groups"] > 1000
    assert abs((E1[0] - E0[0]) / E0[0]) < 1e-9
