"""Benchmark: test-particle steps per second on the 4096-system Planet Nine
sweep (dt = 0.5 yr), one MI355X, through libmint's C ABI.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--form learned|batched|flat|plain]
  python bench.py --workload embryos [--steps K]     (the 1000-embryo all-massive system)
  python bench.py --workload p9_single [--steps K]   (one P9 system with 1e5 test particles)
  python bench.py --workload earth_mars [--steps K]  (Sun + Earth + Mars: energy against the oracle)

A "step" is one Mercurius step of every system of the sweep (about 5.6 M test
particles, five massive bodies each).  The state is uploaded to HBM before the
timed region; the timed region is exactly K steps, bracketed by device
synchronisation (the library's stream event) on both sides.  One JSON line is
printed with the roofline of the dominant kernel (tp_main_kernel, FP64 vector
ALU bound) and the CPU oracle baseline timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "p9-integrator_amd"))

import mint  # noqa: E402
import mint_workloads as W  # noqa: E402

METRIC = "test-particle steps/s, 4096-system P9 sweep, learned+batched vs plain, dE<=1e-9"
# How the line reads the metric's energy gate: the splitting's own |dE/E| at
# dt = 0.5 yr (Jupiter) is ~1e-7, so the gate is the GPU's energy against the
# CPU oracle's after the same steps.
ENERGY_GATE = "|E_gpu - E_oracle| / |E0| <= 1e-9 after the timed steps (massive bodies; the splitting's " \
              "intrinsic |dE/E| at dt 0.5 yr is reported beside it)"
FP64_PEAK = 78.6e12   # MI355X dense FP64 vector, FLOP/s
HBM_PEAK = 8.0e12     # B/s
N_M = 5
# HBM bytes per tp_main_kernel launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (calibrated by a 1 GiB copy: FETCH x2.0, WRITE x1.0), keyed by the launch's
# (block, TPs per lane, steps per launch): {key: (bytes, source)}
MEASURED_TRAFFIC = {(512, 1, 128): (807.5e6, "profiles/tp_main_kernel_p9sweep4096.md (256-step calls, 128-step launches)"),
                    (512, 1, 20): (580.2e6, "profiles/tp_main_kernel_p9sweep4096_20steps.md (20-step calls)"),
                    (256, 1, 128): (807.9e6, "profiles/tp_main_kernel_p9sweep4096.md (256-step calls, block 256)"),
                    (256, 1, 20): (580.9e6, "profiles/tp_main_kernel_p9sweep4096_20steps.md (20-step calls, block 256)")}
PLAIN_BYTES_PER_TP_STEP = 96  # SoA6 state read and written once per step


def flops_per_tp_step(nbar: float, nm: int = N_M) -> float:
    """Algorithmic FLOP per TP-step, as include/mint_flops.h defines it: two
    changeover half-kicks (28 per pair each), two velocity updates (12), two
    jumps (12), the exact encounter check (9 per body), and the
    universal-variable drift (60 + 40 per Newton iteration); add/mul/div/sqrt
    = 1, fma = 2.  The library computes it, so the bench and the header agree."""
    return mint.flops_per_tp_step(nbar, nm)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_ctx(form, guess, scorer, stats=0, **kw):
    g = mint.Mint(dt=0.5, form=form, kepler_guess=guess, use_scorer=scorer, collect_stats=stats, **kw)
    if guess == mint.GUESS_LEARNED or scorer:
        import train_learned as T

        if not os.path.exists(T.DEFAULT_OUT):
            log("bench: no trained weights; training them on this GPU first (untimed)")
            sys.argv = [sys.argv[0]]
            T.main()
        T.apply(g)
    return g


def energy_gate(st, E0, E1, mx, mv, steps):
    """|dE/E| of the massive bodies over the run, and the GPU's energy against
    the CPU oracle's for the same steps (massive bodies only: TPs are
    massless and do not enter the energy)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: checker only

    ms = mint.State(np.zeros(st.n_systems + 1, np.int64), st.m, st.mx, st.mv, np.zeros((0, 3)), np.zeros((0, 3)))
    t0 = time.perf_counter()
    out, _, _ = O.run(ms, steps, 0.5, guess=1, threads=min(16, os.cpu_count() or 1))
    Eo = O.energy(st.m, out["mx"], out["mv"])
    Eg = O.energy(st.m, mx, mv)
    return {"max_rel_dE_over_run": float(np.max(np.abs((E1 - E0) / E0))),
            "max_rel_E_gpu_minus_oracle": float(np.max(np.abs((Eg - Eo) / E0))),
            "massive_state_bitwise_equal_to_oracle": bool(np.array_equal(out["mx"], mx) and
                                                          np.array_equal(out["mv"], mv)),
            "oracle_seconds": time.perf_counter() - t0}


def timed_run(g, st, steps, warmup, repeats=1):
    """Warm up on the uploaded state, build the launch graph of a `steps`-step
    call (host-side capture, so it is not inside the timed call), then time
    `repeats` calls of exactly `steps` steps, each from the same uploaded
    state (re-uploaded into the same buffers; graphs stay valid).  Returns the
    host time of the median call, its event time, the energies before and
    after it, its main-kernel timing, and every call's event time."""
    g.upload(st)
    if warmup > 0:
        g.step(warmup)
    g.prepare(steps)
    runs = []
    for _ in range(max(1, repeats)):
        g.upload(st)
        E0 = g.energy()
        g.reset_counters()
        t0 = time.perf_counter()
        g.step(steps)            # returns after the stream's closing event has completed
        t1 = time.perf_counter()
        tm = g.timing()
        runs.append((tm["ms_total"] / 1e3, t1 - t0, E0, g.energy(), tm))
    order = sorted(range(len(runs)), key=lambda i: runs[i][0])
    ev_s, wall, E0, E1, tm = runs[order[len(runs) // 2]]
    return wall, ev_s, E0, E1, tm, [r[0] * 1e3 for r in runs]


def cpu_baseline(st_full, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: the CPU baseline leg only

    probe = st_full.subset([0])
    t0 = time.perf_counter()
    O.run(probe, 20, 0.5, guess=1)
    rate = probe.n_tp * 20 / max(time.perf_counter() - t0, 1e-6)
    steps = 100
    nsys = int(max(1, min(st_full.n_systems, budget_s * rate / (steps * np.mean(np.diff(st_full.tp_offset))))))
    sample = st_full.subset(np.arange(nsys))
    t0 = time.perf_counter()
    _, _, stats = O.run(sample, steps, 0.5, guess=1, threads=1)
    dt = time.perf_counter() - t0
    # the same sample on every core of this process's share (systems are
    # independent; the oracle splits them over std::threads)
    ncores = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    _, _, stats_n = O.run(sample, steps, 0.5, guess=1, threads=ncores)
    dtn = time.perf_counter() - t0
    return dict(value=stats["tp_steps"] / dt, unit="TP-steps/s", cores=1, kind="port",
                sample=f"systems 0-{nsys - 1} of the sweep ({sample.n_tp} TPs), {steps} steps, oracle single thread, "
                       f"{dt:.1f} s",
                all_cores=dict(value=stats_n["tp_steps"] / dtn, cores=ncores, seconds=round(dtn, 2)),
                cpu=cpu_model())


def cpu_model() -> str:
    """The host CPU's model name and logical CPU count (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            names = [ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs)" if names else "unknown"
    except OSError:
        return "unknown"


def copy_probe(nbytes):
    """One hipMemcpy device-to-device of nbytes (a plain copy kernel of known
    size: FETCH_SIZE / WRITE_SIZE per byte of it calibrate the counters)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def bench_p9_single(args):
    """One P9 system with 1e5 test particles (dt 0.5 yr): the batched form at
    its autotuned shape, the plain form (a launch per step) on fewer steps, and
    the oracle on one core for 100 steps of the same system."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: the CPU baseline leg only

    st = W.p9_single()
    ntp = st.n_tp
    g = make_ctx(mint.FORM_BATCHED, mint.GUESS_SERIES, 0)
    g.upload(st)
    shape = g.autotune(args.tune_cache)
    wall, ev_s, E0, E1, _, _ = timed_run(g, st, args.steps, args.warmup)
    status = int(g.status()[0])
    g.close()
    value = ntp * args.steps / ev_s
    log(f"bench: p9_single batched {shape}: {args.steps} steps in {ev_s * 1e3:.1f} ms -> {value:.4e} TP-steps/s")
    ps = min(args.steps, 1000)
    gpl = make_ctx(mint.FORM_PLAIN, mint.GUESS_CRUDE, 0)
    _, pe, _, _, _, _ = timed_run(gpl, st, ps, 2)
    gpl.close()
    plain = ntp * ps / pe
    t0 = time.perf_counter()
    _, _, so = O.run(st, 100, 0.5, guess=1, threads=1)
    cpu = so["tp_steps"] / (time.perf_counter() - t0)
    out = {"metric": "test-particle steps/s, one P9 system with 1e5 TPs", "value": value, "unit": "TP-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ev_s * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "p9_single_1e5", "tp_total": ntp, "massive": 5, "dt_yr": 0.5,
                      "form": "batched", "launch_shape": shape},
           "forms_tp_steps_per_s": {"batched": value, "plain": plain},
           "cpu_baseline": {"value": cpu, "unit": "TP-steps/s", "cores": 1, "kind": "port",
                            "sample": f"the same system, 100 steps, oracle single thread", "cpu": cpu_model()},
           "energy": {"max_rel_dE_over_run": float(abs((E1[0] - E0[0]) / E0[0])), "status": status}}
    print(json.dumps(out), flush=True)


def bench_earth_mars(args):
    """Sun, Earth and Mars (dt 0.001 yr): the energy check of the plain step,
    the GPU's run against the oracle's over the same steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: checker only

    st = W.earth_mars()
    g = mint.Mint(dt=0.001)
    g.upload(st)
    E0 = g.energy()
    g.reset_counters()
    g.step(args.steps)
    ev_s = g.timing()["ms_total"] / 1e3
    E1 = g.energy()
    fin = g.download()
    g.close()
    t0 = time.perf_counter()
    ref, _, _ = O.run(st, args.steps, 0.001, guess=1)
    cpu_s = time.perf_counter() - t0
    Eo = O.energy(st.m, ref["mx"], ref["mv"])
    out = {"metric": "energy check, Sun + Earth + Mars, dt 0.001 yr", "steps": args.steps,
           "gpu_seconds": ev_s, "oracle_seconds_1_core": cpu_s,
           "max_rel_dE_gpu": float(abs((E1[0] - E0[0]) / E0[0])),
           "max_rel_dE_oracle": float(abs((Eo[0] - E0[0]) / E0[0])),
           "rel_E_gpu_minus_oracle": float(abs((E1[0] - Eo[0]) / E0[0])),
           "massive_state_bitwise_equal_to_oracle": bool(np.array_equal(fin.mx, ref["mx"]) and
                                                         np.array_equal(fin.mv, ref["mv"])),
           "cpu": cpu_model()}
    print(json.dumps(out), flush=True)


EMBRYO_METRIC = "massive body-steps/s, 1000-embryo system (all massive, O(N^2) kicks), dt 0.002"


def flops_per_kick_pass(n: int) -> float:
    """Algorithmic FLOP of one kick-acceleration pass over n massive bodies:
    n (n - 1) ordered pairs at 28 FLOP each (include/mint_flops.h); one pass
    per step, both half-kicks use it."""
    return mint.flops_per_kick_pass(n)


def bench_embryos(args):
    """The all-massive embryo configuration on the many-body engine: K steps
    timed after W warm-up steps, kernel timing from a separate profiling pass,
    the oracle's single-thread rate on the same system as the CPU baseline."""
    st = W.embryos(1000)
    rad = st.meta["radius"]
    n = st.n_massive
    kw = dict(dt=0.002, n_hill=3.0, bs_tol=1e-11)
    g = mint.Mint(**kw)
    g.upload(st)
    g.set_radii(rad)
    if args.warmup > 0:
        g.step(args.warmup)
    g.upload(st)
    g.set_radii(rad)
    E0 = g.energy()
    g.reset_counters()
    t0 = time.perf_counter()
    g.step(args.steps)
    wall = time.perf_counter() - t0
    ev_s = g.timing()["ms_total"] / 1e3
    E1 = g.energy()
    status = int(g.status()[0])
    fin = g.download()
    g.close()
    value = n * args.steps / ev_s
    log(f"bench: embryos: {args.steps} steps in {ev_s * 1e3:.1f} ms -> {value:.4e} body-steps/s")
    # profiling pass: per-kernel HIP-event timing and work counters (no graph)
    gp = mint.Mint(collect_stats=3, **kw)
    gp.upload(st)
    gp.set_radii(rad)
    nprof = min(args.steps, 256)
    gp.reset_counters()
    gp.step(nprof)
    t = gp.timing()
    s = gp.stats()
    gp.close()
    ms = {k: t[f"ms_mb_{k}"] / max(t[f"n_mb_{k}"], 1) for k in ("drift", "close")}
    F = flops_per_kick_pass(n)
    achieved = F / (ms["close"] / 1e3)
    roof = {"bound": "fp64", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK, "traffic": None, "kernel": "mb_close_kernel",
            "flop_per_launch": F, "avg_launch_ms": ms}
    base = None
    check = {}
    if not args.no_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # test infrastructure: checker and CPU baseline leg only

        nb = 1000
        t0 = time.perf_counter()
        ref, _, _ = O.run(st, nb, 0.002, bs_tol=1e-11, radii=rad)
        dt = time.perf_counter() - t0
        base = dict(value=n * nb / dt, unit="body-steps/s", cores=1, kind="port",
                    sample=f"the same 1000-embryo system, {nb} steps, oracle single thread, {dt:.1f} s")
        gc = mint.Mint(**kw)
        gc.upload(st)
        gc.set_radii(rad)
        gc.step(nb)
        out = gc.download()
        gc.close()
        check = {"first_steps_bitwise_equal_to_oracle": bool(np.array_equal(out.mx, ref["mx"]) and
                                                              np.array_equal(out.mv, ref["mv"])),
                 "steps_compared": nb}
        log(f"bench: cpu baseline {base}")
    out = {
        "metric": EMBRYO_METRIC, "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ev_s * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "embryos_1000", "bodies": n, "dt_yr": 0.002, "n_hill": 3.0, "bs_tol": 1e-11,
                   "engine": "many-body"},
        "roofline": roof, "cpu_baseline": base,
        "energy": {"max_rel_dE_over_run": float(abs((E1[0] - E0[0]) / E0[0])), "status_bits": status, **check},
        "profile": {"steps": nprof, "stats": s, "timing": t,
                    "groups_per_step": s["massive_groups"] / nprof,
                    "bodies_in_groups_per_step": s["massive_group_bodies"] / nprof},
        "host_wall_s": wall,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default 2000 (sweep), 10000 (embryos)")
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--repeats", type=int, default=5, help="timed calls of --steps steps; the median is reported")
    ap.add_argument("--form", default="learned", choices=["learned", "batched", "flat", "plain"])
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--no-baseline", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--no-ablation", action="store_true", help="skip the plain / flat comparison runs")
    ap.add_argument("--plain-systems", type=int, default=64, help="systems in the plain-form sample")
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "p9-integrator_amd", "autotune.cache"))
    ap.add_argument("--shape", default=None, help="block,tps_per_lane,window: skip the autotuner")
    ap.add_argument("--workload", default="p9_sweep", choices=["p9_sweep", "embryos", "p9_single", "earth_mars"],
                    help="embryos: the 1000-embryo all-massive system (default steps 10000); p9_single: one "
                         "P9 system with 1e5 TPs (10000); earth_mars: the energy check (100000)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus must be 1 (independent systems would shard "
                                   "across GPUs by system range, with no exchange between them)"}))
        sys.exit(2)
    if args.steps is None:
        args.steps = {"embryos": 10000, "p9_single": 10000, "earth_mars": 100000}.get(args.workload, 2000)
    if args.workload == "embryos":
        bench_embryos(args)
        return
    if args.workload == "p9_single":
        bench_p9_single(args)
        return
    if args.workload == "earth_mars":
        bench_earth_mars(args)
        return

    log(f"bench: generating the {args.systems}-system sweep")
    st = W.p9_sweep(S=4096, systems=np.arange(args.systems))
    ntp = st.n_tp
    log(f"bench: {ntp} TPs")

    form = {"learned": mint.FORM_BATCHED, "batched": mint.FORM_BATCHED, "flat": mint.FORM_FLAT,
            "plain": mint.FORM_PLAIN}[args.form]
    guess = mint.GUESS_LEARNED if args.form == "learned" else mint.GUESS_SERIES
    scorer = 1 if args.form == "learned" else 0

    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else None
    # collect_stats bit 2: each main-kernel launch of the timed call is timed
    # by events around it on its own stream, inside the replayed graph
    kt = 4 if form == mint.FORM_BATCHED else 0
    if shape:
        g = make_ctx(form, guess, scorer, stats=kt, block_size=shape[0], tps_per_lane=shape[1],
                     chunk_steps=shape[2])
    elif scorer:
        # the scorer is trained for one launch window: tune block and TPs per lane at that window
        import train_learned as T

        g = make_ctx(form, guess, scorer, stats=kt, chunk_steps=T.window())
    else:
        g = make_ctx(form, guess, scorer, stats=kt)
    tuned = None
    if form == mint.FORM_BATCHED and not shape:
        g.upload(st)
        tuned = tuple(g.autotune(args.tune_cache))
        # the shape a call of --steps steps uses (calls of one launch window
        # have their own tuned block and TPs per lane)
        shape = tuple(g.launch_shape(args.steps))
        log(f"bench: launch shape block={shape[0]} tps/lane={shape[1]} window={shape[2]} for {args.steps}-step "
            f"calls (longer calls: {tuned})")
    wall, ev_s, E0, E1, tm, calls_ms = timed_run(g, st, args.steps, args.warmup, args.repeats)
    status = g.status()
    fin = g.download()
    g.close()
    value = ntp * args.steps / ev_s
    log(f"bench: {args.form}: {args.steps} steps in {ev_s * 1e3:.3f} ms (events, median of {len(calls_ms)} calls "
        f"{[round(x, 3) for x in calls_ms]}) / {wall * 1e3:.3f} ms (host) -> {value:.4e} TP-steps/s")
    dE = float(np.max(np.abs((E1 - E0) / E0)))

    extra = {}
    # n-bar as the algorithmic work per TP-step defines it: the mean Newton
    # iterations of the plain form's drift (crude guess s0 = dt/r0) on this
    # input, counted on the GPU over the first steps of every system (the same
    # kernel and guess as the plain form, all systems in one launch)
    gn = make_ctx(mint.FORM_FLAT, mint.GUESS_CRUDE, 0, stats=1)
    gn.upload(st)
    nbar_steps = min(args.steps, 20)
    gn.step(nbar_steps)
    sn = gn.stats()
    gn.close()
    nbar = sn["newton_iters"] / max(sn["kepler_solves"], 1)
    F = flops_per_tp_step(nbar)

    roof = None
    if form == mint.FORM_BATCHED:
        # work counts of the same call (a replay with the counters on: the
        # work is a deterministic function of the input)
        gp = make_ctx(form, guess, scorer, stats=1, block_size=shape[0], tps_per_lane=shape[1], chunk_steps=shape[2])
        gp.upload(st)
        gp.reset_counters()
        gp.step(args.steps)
        s = gp.stats()
        gp.close()
        main_steps = s["tp_steps"] - s["enc_tp_steps"]
        n_main = max(tm["n_tp_main"], 1)
        avg_main_ms = tm["ms_tp_main"] / n_main
        kernel_ms_per_step = tm["ms_tp_main"] / args.steps
        ok = tm["n_tp_main"] > 0 and tm["ms_tp_main"] <= ev_s * 1e3
        achieved = F * main_steps / (tm["ms_tp_main"] / 1e3) if ok else None
        C_eff = min(shape[2], args.steps)
        traffic = MEASURED_TRAFFIC.get((shape[0], shape[1], C_eff))
        roof = {"bound": "fp64", "achieved": achieved / 1e12 if ok else None, "peak": FP64_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK if ok else None,
                "traffic": traffic[0] if traffic else None, "traffic_source": traffic[1] if traffic else None,
                "kernel": "tp_main_kernel", "flop_per_tp_step": F, "newton_mean_plain_form": nbar,
                "newton_mean_this_form": s["newton_iters"] / max(s["kepler_solves"], 1),
                "tp_steps_per_launch": main_steps / n_main, "avg_launch_ms": avg_main_ms,
                "launches": tm["n_tp_main"], "kernel_ms_per_step": kernel_ms_per_step,
                "timing": "HIP events around each tp_main_kernel launch of the timed call, on its stream"}
        if not ok:
            roof["error"] = "main-kernel time missing or longer than the timed call; frac withheld"
        extra["profile"] = {"steps": args.steps, "stats": s, "bs_fraction": s["bs_steps"] / max(s["tp_steps"], 1)}

    forms = {args.form: value}
    if not args.no_ablation and args.form == "learned":
        gb = make_ctx(mint.FORM_BATCHED, mint.GUESS_SERIES, 0, block_size=shape[0], tps_per_lane=shape[1],
                      chunk_steps=shape[2])
        _, be, _, _, _, _ = timed_run(gb, st, args.steps, args.warmup, args.repeats)
        forms["batched"] = ntp * args.steps / be
        gb.close()
    if not args.no_ablation and form == mint.FORM_BATCHED:
        # flat: one launch per step over all systems
        gf = make_ctx(mint.FORM_FLAT, mint.GUESS_SERIES, 0)
        fs = min(args.steps, 200)
        _, fe, _, _, _, _ = timed_run(gf, st, fs, 2)
        forms["flat"] = ntp * fs / fe
        gf.close()
        # plain: the drop-in used like the reference uses its kernel (4 systems
        # per launch, a launch per step, crude Kepler guess); timed on a sample
        ps = st.subset(np.arange(min(args.plain_systems, st.n_systems)))
        gpl = make_ctx(mint.FORM_PLAIN, mint.GUESS_CRUDE, 0)
        _, pe, _, _, _, _ = timed_run(gpl, ps, args.plain_steps, 2)
        forms["plain"] = ps.n_tp * args.plain_steps / pe
        extra["plain_sample"] = f"systems 0-{ps.n_systems - 1} ({ps.n_tp} TPs), {args.plain_steps} steps"
        gpl.close()
        log(f"bench: forms {forms}")
        # HBM roofline of the per-step kernel (96 B per TP-step), from HIP
        # events around each launch: the plain form's launches of 4 systems,
        # and the flat form's launches over all systems
        per_step = {}
        for name, f_, sample, nst in (("plain", mint.FORM_PLAIN, ps, 20), ("flat", mint.FORM_FLAT, st, 10)):
            gt = make_ctx(f_, mint.GUESS_CRUDE, 0, stats=2)
            gt.upload(sample)
            gt.step(2)
            gt.reset_counters()
            gt.step(nst)
            t = gt.timing()
            gt.close()
            nl = max(t["n_tp_main"], 1)
            ms = t["ms_tp_main"] / nl
            tps = sample.n_tp * nst / nl
            ach = PLAIN_BYTES_PER_TP_STEP * tps / (ms / 1e3)
            per_step[name] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                              "frac": ach / HBM_PEAK, "traffic": None, "kernel": "tp_step_kernel",
                              "bytes_per_tp_step": PLAIN_BYTES_PER_TP_STEP, "tp_steps_per_launch": tps,
                              "avg_launch_ms": ms, "launches": t["n_tp_main"]}
        extra["roofline_per_step_forms"] = per_step

    base = None if args.no_baseline else cpu_baseline(st)
    if base:
        log(f"bench: cpu baseline {base}")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TP-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ev_s * 1e3 / args.steps,
        "timed_calls_ms": calls_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "p9_sweep_4096", "systems": st.n_systems, "tp_total": ntp, "massive_per_system": 5,
                   "dt_yr": 0.5, "form": args.form, "kepler_guess": int(guess), "scorer": scorer,
                   "launch_shape": list(shape) if shape else None,
                   "launch_shape_multiwindow_calls": list(tuned) if tuned else None},
        "roofline": roof,
        "cpu_baseline": base,
        "forms_tp_steps_per_s": forms,
        "speedup_vs_plain": value / forms["plain"] if "plain" in forms else None,
        "energy": dict(energy_gate(st, E0, E1, fin.mx, fin.mv, args.steps) if not args.no_baseline else
                       {"max_rel_dE_over_run": dE}, status_bits_or=int(np.bitwise_or.reduce(status))),
        "energy_gate": ENERGY_GATE,
        "host_wall_s": wall,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
