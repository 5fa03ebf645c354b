"""Massive encounter groups and the many-body engine on the GPU, bit for bit
against the oracle: the 1000-embryo configuration, a dense embryo ring with
multi-body groups and physical radii, the lane-per-system engine and the
many-body engine on the same 8-body set, sixteen 8-body clusters eight to a
wavefront, a co-orbital pair beside test particles, graph replay and call
splitting, and checkpoint/resume."""
import numpy as np
import pytest

import mint
import mint_workloads as W
import oracle as O
from gpu_cases import coorbital_system, embryo_cluster, embryo_ring

pytestmark = pytest.mark.gpu


def _gpu(st, nsteps, dt, radii=None, calls=1, **kw):
    kw.setdefault("collect_stats", 1)
    with mint.Mint(dt=dt, **kw) as g:
        g.upload(st)
        if radii is not None:
            g.set_radii(radii)
        per = [nsteps // calls + (1 if c < nsteps % calls else 0) for c in range(calls)]
        for n in per:
            if n:
                g.step(n)
        return g.download(), g.status(), g.stats(), g.energy()


def _same(out, ref):
    assert np.array_equal(out.mx, ref["mx"]) and np.array_equal(out.mv, ref["mv"])
    assert np.array_equal(out.tx, ref["tx"]) and np.array_equal(out.tv, ref["tv"])


def test_embryos_1000_bitwise_20_steps():
    st = W.embryos(1000)
    rad = st.meta["radius"]
    ref, rst, rstats = O.run(st, 20, 0.002, bs_tol=1e-11, radii=rad)
    out, status, stats, _ = _gpu(st, 20, 0.002, radii=rad, bs_tol=1e-11)
    _same(out, ref)
    assert np.array_equal(status, rst)
    for k in ("massive_groups", "massive_group_bodies", "overlaps"):
        assert stats[k] == rstats[k], k
    assert rstats["massive_groups"] > 0


def test_embryos_at_the_body_limit_bitwise():
    """2048 bodies, the most one system takes: the largest staged windows and
    chunk-sum buffers of the pair-check and kick kernels."""
    st = W.embryos(2048)
    rad = st.meta["radius"]
    ref, rst, rstats = O.run(st, 5, 0.002, bs_tol=1e-11, radii=rad)
    out, status, stats, _ = _gpu(st, 5, 0.002, radii=rad, bs_tol=1e-11)
    _same(out, ref)
    assert np.array_equal(status, rst)
    for k in ("massive_groups", "massive_group_bodies", "overlaps"):
        assert stats[k] == rstats[k], k


@pytest.mark.parametrize("n,steps", [(64, 300), (200, 150)])
def test_embryo_ring_bitwise(n, steps):
    st = embryo_ring(n)
    ref, rst, rstats = O.run(st, steps, 0.002, bs_tol=1e-11, radii=st.meta["radius"])
    out, status, stats, E = _gpu(st, steps, 0.002, radii=st.meta["radius"], bs_tol=1e-11)
    _same(out, ref)
    assert np.array_equal(status, rst)
    assert stats["massive_groups"] == rstats["massive_groups"] > 0
    assert stats["massive_group_bodies"] == rstats["massive_group_bodies"]
    assert rstats["massive_group_bodies"] > 2 * rstats["massive_groups"]  # some groups of 3+
    assert np.array_equal(E, O.energy(st.m, ref["mx"], ref["mv"]))


@pytest.mark.parametrize("n,span,steps,mean_size", [(30, 0.3, 40, 4), (48, 0.25, 40, 15), (12, 0.02, 100, 12),
                                                    (24, 0.06, 100, 24), (40, 0.08, 20, 40), (64, 0.1, 20, 64)])
def test_every_group_lane_layout_bitwise(n, span, steps, mean_size):
    """Encounter groups in every lane layout of the wavefront group BS: groups
    of 2..8 members (all eight Gragg rows side by side), 12 (rows four at a
    time), 24 (two at a time), 40 and 64 members (one row at a time), and a
    mix of sizes; bit for bit against the oracle."""
    st = embryo_cluster(n, span=span)
    print(f"embryo cluster n={n} span={span} steps={steps}", flush=True)
    ref, rst, rstats = O.run(st, steps, 0.002, bs_tol=1e-11, radii=st.meta["radius"])
    out, status, stats, _ = _gpu(st, steps, 0.002, radii=st.meta["radius"], bs_tol=1e-11)
    assert rstats["massive_group_bodies"] >= mean_size * rstats["massive_groups"] > 0
    _same(out, ref)
    assert np.array_equal(status, rst)
    for k in ("massive_groups", "massive_group_bodies", "overlaps"):
        assert stats[k] == rstats[k], k


def test_many_body_systems_side_by_side():
    """Three 40-body systems in one call (one workgroup row per system): a
    ring with encounter groups, a cluster with groups and overlaps, and a
    quiet ring.  Every system ends bit for bit where the oracle puts it, with
    its own status bits, and the middle one as if it had run alone."""
    parts = [embryo_ring(40, seed=1), embryo_cluster(40, span=0.25, seed=2),
             embryo_ring(40, seed=3, a_lo=1.5, a_hi=1.53)]
    rad = np.concatenate([p.meta["radius"] for p in parts])
    st = mint.State(np.zeros(4, np.int64), np.concatenate([p.m for p in parts]),
                    np.concatenate([p.mx for p in parts]), np.concatenate([p.mv for p in parts]),
                    np.zeros((0, 3)), np.zeros((0, 3)))
    ref, rst, rstats = O.run(st, 100, 0.002, bs_tol=1e-11, radii=rad)
    out, status, stats, _ = _gpu(st, 100, 0.002, radii=rad, bs_tol=1e-11)
    _same(out, ref)
    assert np.array_equal(status, rst) and len(set(rst.tolist())) == 3
    for k in ("massive_groups", "massive_group_bodies", "overlaps"):
        assert stats[k] == rstats[k], k
    alone, _, _ = O.run(parts[1], 100, 0.002, bs_tol=1e-11, radii=parts[1].meta["radius"])
    assert np.array_equal(out.mx[1], alone["mx"][0]) and np.array_equal(out.mv[1], alone["mv"][0])


def test_lane_engine_and_many_body_engine_agree_on_8_bodies():
    st = embryo_cluster(8, seed=0)
    ref, rst, rstats = O.run(st, 400, 0.002, bs_tol=1e-11)
    assert rstats["massive_group_bodies"] > 2 * rstats["massive_groups"] > 0
    for many in (0, 1):
        for form in (mint.FORM_PLAIN, mint.FORM_BATCHED):
            out, status, stats, _ = _gpu(st, 400, 0.002, bs_tol=1e-11, many_body=many, form=form)
            _same(out, ref)
            assert np.array_equal(status, rst)
            assert stats["massive_groups"] == rstats["massive_groups"], (many, form)


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_BATCHED])
def test_lane_engine_groups_in_every_system_of_a_wave(form):
    """Sixteen 8-body clusters in the lane-per-body engine: eight systems to a
    wavefront, most of them inside a massive encounter group at the same
    step, each running its group BS from its own workspace.  Every system
    ends bit for bit where the oracle puts it."""
    parts = [embryo_cluster(8, seed=10 + k) for k in range(16)]
    st = mint.State(np.zeros(17, np.int64), np.concatenate([p.m for p in parts]),
                    np.concatenate([p.mx for p in parts]), np.concatenate([p.mv for p in parts]),
                    np.zeros((0, 3)), np.zeros((0, 3)))
    ref, rst, rstats = O.run(st, 200, 0.002, bs_tol=1e-11)
    assert rstats["massive_groups"] > 8 * 200 and rstats["massive_group_bodies"] > 2 * rstats["massive_groups"]
    out, status, stats, _ = _gpu(st, 200, 0.002, bs_tol=1e-11, many_body=0, form=form)
    _same(out, ref)
    assert np.array_equal(status, rst)
    for k in ("massive_groups", "massive_group_bodies", "overlaps"):
        assert stats[k] == rstats[k], k


@pytest.mark.parametrize("form", [mint.FORM_PLAIN, mint.FORM_FLAT, mint.FORM_BATCHED])
def test_coorbital_group_beside_test_particles(form):
    st = coorbital_system()
    ref, rst, rstats = O.run(st, 40, 0.5)
    assert rstats["massive_groups"] == 40
    out, status, stats, _ = _gpu(st, 40, 0.5, form=form)
    _same(out, ref)
    assert np.array_equal(status, rst)
    assert stats["massive_groups"] == rstats["massive_groups"]


def test_many_body_graph_and_call_split_are_bitwise():
    st = embryo_ring(100, seed=4)
    ref, _, _ = O.run(st, 150, 0.002, bs_tol=1e-11)
    for graph, calls in ((1, 1), (0, 1), (1, 3)):
        out, _, _, _ = _gpu(st, 150, 0.002, calls=calls, bs_tol=1e-11, use_graph=graph)
        _same(out, ref)


def test_many_body_checkpoint_resume(tmp_path):
    st = embryo_ring(100, seed=6)
    rad = st.meta["radius"]
    path = str(tmp_path / "emb.ckpt")
    full, _, _, _ = _gpu(st, 90, 0.002, radii=rad, bs_tol=1e-11)
    with mint.Mint(dt=0.002, bs_tol=1e-11, collect_stats=1) as g:
        g.upload(st)
        g.set_radii(rad)
        g.step(50)
        g.save_checkpoint(path, 50)
    with mint.Mint(dt=0.002, bs_tol=1e-11, collect_stats=1) as g:
        assert g.load_checkpoint(path) == 50
        g.step(40)
        out = g.download()
    assert np.array_equal(out.mx, full.mx) and np.array_equal(out.mv, full.mv)


def test_many_body_refuses_test_particles():
    st = embryo_ring(12)
    st.tp_offset = np.array([0, 1], np.int64)
    st.tx = np.array([[2.0, 0, 0]])
    st.tv = np.array([[0, 4.4, 0]])
    with mint.Mint(dt=0.002) as g:
        with pytest.raises(mint.MintError, match="test particles"):
            g.upload(st)
