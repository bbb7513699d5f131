"""Benchmark of the random-walk hot path on one MI355X.

Headline workload (the precision workload): the heterogeneous 200x200x20
lognormal field (1 m cells, divergence-free face flows, porosity
U[0.20,0.30], aL=0.5 m, aT=0.05 m, Dm=1e-4 m2/d) with 1e7 particles
released uniformly in the 10x50x20-cell inlet block, FIXED dt = 0.5 d.  One
bench step is one tracking interval: one rwpt_advance of every particle by
50 RW time steps (25 d), then one gpkde reconstruction (one density pass) of
the active particles on the 200x200x20 grid, as MPathRW does at a gpkde
time point.  The run is 20 intervals (1000 RW steps); --steps beyond 20
starts the run again from the release (the reload from host memory is then
inside the timed region).  Particle state is resident in HBM before the
timed region; --warmup intervals run first and the state is then reloaded.

value = accepted RW sub-steps (the iterations of the reference's
TrackSubCell loop) per second of the timed region (walk and
reconstructions), FP32-local walk (--precision 1: FP32 cell-local positions
and sub-step arithmetic, FP64 times, the FP64 walk's random draws).  The
plain FP64 walk runs the same timed workload for the speed-up; an untimed
pass of both walks gives the FP32 errors on the plume mean and variance and
on the concentration at 64 observation cells.  When a moment error exceeds
0.5 % the value is the FP64 rate instead (error_gate).

The roofline is that of the walk kernel over the timed region (HIP events
on the engine stream): algorithmic bytes = the SoA state one particle-step
reads and writes (96 B FP64, 72 B FP32-local) times the launch's
particle-steps, HBM bound, 8 TB/s; traffic = the measured HBM bytes per
launch of the committed rocprofv3 summary of this workload.  cpu_baseline =
the reference's own tracking engine compiled from its sources (OpenMP, one
rng stream per thread) on the first interval of 180,000 of the particles,
beside the oracle (OpenMP) on 1e6.

Further legs of the same line: config1 (1e6 particles, 10 intervals, FP32
and FP64 walk rates), kde_adaptive (gpkde bandwidth optimisation of the
plume after one interval, beside one loop of the compiled reference),
ensemble (256 members, about 6.3e7 particles, one walk over all members)
and adjoint (the calibration workload: 18 zoned parameters, the pathwise
gradient of an observation misfit, its cost and its check against central
differences).  --quick leaves them out.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modpath-rw_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# algorithmic bytes of one particle-step: the state a step reads and writes
# (local x y z, cell + status, 2 x u64 rng), read + write
BYTES_PER_STEP = {0: 2 * (24 + 8 + 16), 1: 2 * (12 + 8 + 16), 2: 2 * (12 + 8 + 16)}
HBM_PEAK_GBS = 8000.0
INTERVAL_STEPS = 50
RUN_INTERVALS = 20          # 1000 RW steps
DT = 0.5
HEADLINE_PARTICLES = 10_000_000
ERROR_BOUND = 0.005
PRECISION_NAME = {0: "plain FP64", 1: "FP32 cell-local, FP64 time, the FP64 walk's draws",
                  2: "FP32 cell-local, FP64 time, 4-draw normals"}
TRAFFIC_FILE = {0: "walk_kernel_fp64_config2_traffic.txt",
                1: "walk_kernel_fp32_config2_traffic.txt",
                2: "walk_kernel_fp32n4_config2_traffic.txt"}
TUNED_NAME = {0: "config2_fp64", 1: "config2_fp32", 2: "config2_fp32n4"}
TUNED_NAME_1 = {0: "config1_fp64", 1: "config1_fp32", 2: "config1_fp32n4"}


def workload(n_particles):
    """The config-1 field and n_particles released uniformly in the inlet
    block (columns 1-10, rows 76-125, all 20 layers)."""
    from rwpt import problems
    P = problems.config1()
    rel = problems.release_in_block(P, n_particles, (1, 10), (76, 125), (1, 20))
    return P, rel


def cpu_baseline(P, rel, n_sample, n_reference=180000):
    """The host cores on the first tracking interval of the same workload:
    the reference's own tracking engine compiled from its sources
    (oracle/_ref/rwpt_ref in timing mode: OpenMP over particles with one rng
    stream per thread, as MPathRW runs it) on n_reference particles when it
    is built, and the oracle (plain C, OpenMP over particles) on n_sample
    particles.  The reference reports no sub-step count, so its rate is the
    oracle's sub-step count on the same particles over the reference's
    tracking time (the two walks draw different numbers, so the counts agree
    statistically, not exactly)."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1

    def port(n):
        cell, lx, ly, lz = (a[:n] for a in rel)
        st = O.new_state(cell, lx, ly, lz)
        t0 = time.perf_counter()
        O.advance(O.OracleProblem(P), st, 0.0, INTERVAL_STEPS * DT, nthreads=threads)
        return int(st["substeps"].sum()), time.perf_counter() - t0

    steps, el = port(n_sample)
    out = {"value": steps / el, "unit": "particle-steps/s", "cores": threads,
           "kind": "port",
           "sample": f"{n_sample} particles x {INTERVAL_STEPS} RW steps "
                     f"({steps} particle-steps) of the same workload, "
                     f"{el:.1f} s on {threads} OpenMP threads"}
    try:
        from oracle import refrun
        if not refrun.available():
            return out
        n = min(n_reference, n_sample)
        cell, lx, ly, lz = (a[:n] for a in rel)
        tm = {}
        refrun.run(P, cell, lx, ly, lz, [INTERVAL_STEPS * DT], threads=threads, timing=tm,
                   timeout=300)
        rsteps, _ = port(n)
        port_out = dict(out)
        out = {"value": rsteps / tm["seconds"], "unit": "particle-steps/s",
               "cores": threads, "kind": "reference",
               "sample": f"{n} particles x {INTERVAL_STEPS} RW steps of the same workload, "
                         f"{tm['seconds']:.1f} s of tracking on {threads} OpenMP threads; "
                         f"particle-steps = the oracle's {rsteps} sub-steps on the same "
                         "particles (the reference counts none)",
               "port": port_out}
    except Exception as exc:  # the port's figure stands without the reference
        out["reference_error"] = str(exc)
    return out


def adaptive_kde(E, origin, nbins, loops):
    """gpkde with the bandwidth optimisation (brute-force kernels,
    convergence tests on, as the reference's example input sets it) on the
    engine's active particles; the compiled reference, when it is built,
    runs one loop of the same reconstruction on the host cores."""
    cells = int(np.prod(nbins))
    E.kde(origin, (1.0, 1.0, 1.0), nbins, mode=1, n_opt_loops=1)  # warm
    k0, _ = E.kernel_time(2)
    t0 = time.perf_counter()
    _, hist = E.kde(origin, (1.0, 1.0, 1.0), nbins, mode=1, n_opt_loops=loops)
    wall = time.perf_counter() - t0
    k1, _ = E.kernel_time(2)
    run = E.kde_info()["loops"]
    out = {"unit": "KDE cells/s per optimisation loop", "loops_requested": loops,
           "loops_run": run, "active_bins": int((hist != 0).sum()),
           "wall_ms": wall * 1e3, "device_ms": k1 - k0,
           "ms_per_loop": wall * 1e3 / max(run, 1),
           "cells_per_s": cells * max(run, 1) / wall}
    try:
        from oracle import refrun
        if refrun.gpkde_available():
            g = E.particles()
            act = g["status"] == 1
            gx, gy, gz = E.global_xyz()
            ts = []
            for L in (0, 1):
                t = time.perf_counter()
                refrun.gpkde(gx[act], gy[act], gz[act], origin, (1.0, 1.0, 1.0), nbins,
                             n_opt_loops=L, unit_volume=False, skip_error=True,
                             threads=os.cpu_count() or 1)
                ts.append(time.perf_counter() - t)
            out["cpu_reference"] = {
                "ms_per_loop": (ts[1] - ts[0]) * 1e3, "ms_without_loops": ts[0] * 1e3,
                "cores": os.cpu_count() or 1, "kind": "reference",
                "sample": f"the same {int(act.sum())} points, one optimisation loop, "
                          "the reference's GridProjectedKDE compiled from its "
                          "sources (amdflang -O2 -fopenmp)"}
    except Exception as exc:  # the GPU figure stands without the CPU one
        out["cpu_reference"] = {"error": str(exc)}
    return out




def observation_cells(P):
    """64 observation cells (0-based flat indices, x fastest) along the
    plume's path: 8 columns x 8 rows in the middle layer."""
    cols = np.linspace(15, 150, 8).astype(int)
    rows = np.linspace(80, 120, 8).astype(int)
    k = P.nlay // 2
    return np.array([c + P.ncol * (r + P.nrow * k) for r in rows for c in cols])


def apply_tuning(E, name, fallback=None):
    """The walk's launch shape for this workload as the autotuner
    (python -m rwpt.autotune, rwpt_autotune) stored it; the fallback
    workload's shape, else the defaults, when it has no entry."""
    from rwpt import autotune
    for key in (name, fallback):
        t = autotune.tuned(key) if key else None
        if t is not None:
            E.set_tuning(budget=t["budget"], block=t["block"], sort=bool(t["sort"]),
                         min_waves=t["min_waves"])
            return dict(t, source="modpath-rw_amd/rwpt/tuning.json (rwpt_autotune)",
                        workload=key)
    return {"source": "defaults"}


def ensemble_run(intervals, warmup):
    """The batched workload: 256 ensemble members on one 100x100x10 field,
    per-member dispersivities and porosity, log-uniform particle counts
    (1e4..1e6, about 5.5e7 in all), FP32-local, all members in the same
    walk launches; `intervals` tracking intervals of 50 RW steps timed."""
    from rwpt import Engine, problems
    E, counts = problems.config3()
    rel = problems.release_ensemble(E, counts, (1, 5), (38, 63), (1, 10))
    eng = Engine(E, precision=1)
    tune = apply_tuning(eng, "ensemble256_fp32")
    try:
        eng.load_particles(*rel)
        eng.set_time(0.0)
        for w in range(warmup):
            eng.advance((w + 1) * INTERVAL_STEPS * DT)
        eng.load_particles(*rel)
        eng.set_time(0.0)
        eng.reset_kernel_time()
        steps = 0
        t0 = time.perf_counter()
        for k in range(intervals):
            steps += eng.advance((k + 1) * INTERVAL_STEPS * DT)[1]
        el = time.perf_counter() - t0
        kms, launches = eng.kernel_time(0)
        active = int((eng.particles()["status"] == 1).sum())
    finally:
        eng.close()
    return {"unit": "particle-steps/s", "value": steps / el,
            "workload": "256 members on one seeded 100x100x10 lognormal field, "
                        "aL ~ U[0.05,1], aT = aL*U[0.05,0.2], porosity ~ U[0.2,0.35] "
                        "per member, particles log-uniform in [1e4, 1e6] per member, "
                        "FP32-local, FIXED dt 0.5 d",
            "members": int(E.members), "particles": int(counts.sum()),
            "intervals": intervals, "rw_steps_per_interval": INTERVAL_STEPS,
            "particle_steps": steps, "ms_per_interval": el * 1e3 / intervals,
            "walk_launches": launches,
            "walk_kernel_avg_launch_ms": kms / max(launches, 1),
            "active_after": active, "tuning": tune}


def adjoint_run(n=1_000_000, seeds=8, rel_steps=(1e-4, 1e-2)):
    """The calibration workload: the config-1 field with 8 x-bands of 25
    columns (own aL, aT each), a common Dm and porosity scale (18
    parameters), 1e6 particles, 300 RW steps of 0.25 d in 12 checkpoint
    windows of 25 steps, misfit of the Gaussian-smoothed (h = 2, 2, 1 m)
    concentrations at 32 observation points over the plume against 0.8 x
    the model's own.

    Times the forward (rwpt_adjoint without gradient) and the adjoint, and
    checks the gradient with respect to aL of zone 1 and to the porosity
    scale:
    - on a 4-step horizon against central differences at the largest
      relative step of 1e-8, 1e-9, 1e-10 at which no particle changes
      branch between the +- runs (the two must then agree to rounding);
    - on the full horizon against central differences at the relative
      steps `rel_steps` with common random numbers within each +- pair,
      averaged over `seeds` base seeds, beside the gradient estimators of
      the same seeds: the exact pathwise gradient, the adjoint restarted
      straight-through at every window start (tangent_windows 1), the
      gradient without the random displacements' derivative with respect
      to the sub-step duration (frozen_noise), and both together.
      Reports each estimator's seed mean, standard error and bias against
      the central differences at each step (paired by seed, with the
      standard error of the difference);
    - with the sweep's diagnostics (seed 0, pathwise): the adjoint growth
      per iteration kind (sub-step, jump-limiter cut, exit, grazing exit,
      rare), the exits by their distance from a double root, and each
      particle's growth over the horizon."""
    from rwpt import Engine, problems
    from rwpt.engine import default_options
    dt, every, total = 0.25, 25, 300
    P, zones, (aL, aT, _, dm) = problems.config4(ts_params=(dt, 0.0))
    rel = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    bw = (2.0, 2.0, 1.0)
    seed0 = tuple(int(x) for x in default_options().base_seed)

    def run(windows, obs, a_l=aL, por=1.0, grad=False, target=None, seed=0, opts=None,
            state=False):
        Q = problems.zoned_dispersion(P, zones, a_l, aT, dm)
        Q.porosity = P.porosity * por
        E = Engine(Q, base_seed=((seed0[0] + 7919 * seed) % 2**64, seed0[1]))
        try:
            apply_tuning(E, TUNED_NAME_1[0])
            if opts:
                E.adjoint_options(**opts)
            E.load_particles(*rel)
            out = E.adjoint(windows, obs, np.zeros(len(obs)) if target is None else target,
                            bw, gradient=grad)
            if opts and opts.get("diagnostics"):
                out["stats"] = E.adjoint_stats()
            if state:
                g = E.particles()
                out["state"] = (g["cell"].copy(), g["status"].copy(), g["substeps"].copy())
        finally:
            E.close()
        return out

    def plume_points(windows):
        """8 x 4 observation points over the active plume after the windows
        (mean +- 1.5 standard deviations in x and y, middle layer)."""
        Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
        E = Engine(Q)
        try:
            E.load_particles(*rel)
            for t in windows:
                E.advance(t)
            on = E.particles()["status"] == 1
            gx, gy, _ = E.global_xyz()
        finally:
            E.close()
        mx, my, sx, sy = gx[on].mean(), gy[on].mean(), gx[on].std(), gy[on].std()
        k = P.nlay // 2
        zmid = 0.5 * (P.top[k * P.nrow * P.ncol] + P.bot[k * P.nrow * P.ncol])
        return np.array([(x, y, zmid) for y in np.linspace(my - 1.5 * sy, my + 1.5 * sy, 4)
                         for x in np.linspace(mx - 1.5 * sx, mx + 1.5 * sx, 8)])

    def central(windows, obs, tgt, param, h, seed, changes=False):
        vals = []
        for sg in (1, -1):
            if param == "aL_zone1":
                a2 = aL.copy()
                a2[0] *= 1 + sg * h
                vals.append(run(windows, obs, a_l=a2, target=tgt, seed=seed, state=changes))
            else:
                vals.append(run(windows, obs, por=1 + sg * h, target=tgt, seed=seed,
                                state=changes))
        fd = (vals[0]["misfit"] - vals[1]["misfit"]) / (
            2 * h * (aL[0] if param == "aL_zone1" else 1.0))
        if not changes:
            return fd
        moved = sum(int((a != b).sum()) for a, b in zip(vals[0]["state"], vals[1]["state"]))
        return fd, moved

    def pick(out):
        return {"aL_zone1": float(out["grad_al"][zones == 0].sum()),
                "porosity_scale": float(out["grad_porosity_scale"])}

    params = ("aL_zone1", "porosity_scale")
    # 4-step check: no branch changes at 1e-8
    short = [dt * 4]
    obs_s = plume_points(short)
    tgt_s = 0.8 * run(short, obs_s)["conc"]
    g_short = pick(run(short, obs_s, grad=True, target=tgt_s))
    # the largest of these steps at which no particle's cell, status or
    # sub-step count differs between the +- runs (a branch change moves a
    # particle by a fraction of a cell, not by O(h))
    fd_short = {}
    for p in params:
        tried = []
        for h in (1e-8, 1e-9, 1e-10):
            fd, moved = central(short, obs_s, tgt_s, p, h, 0, changes=True)
            tried.append({"rel_step": h, "central_difference": fd,
                          "state_entries_changed": moved})
            if moved == 0:
                break
        fd_short[p] = tried

    # full horizon: cost, diagnostics, seed-averaged central differences
    full = [dt * every * (k + 1) for k in range(total // every)]
    obs = plume_points(full)
    tgt = 0.8 * run(full, obs)["conc"]
    fwd = run(full, obs, target=tgt)
    A = run(full, obs, grad=True, target=tgt)
    diag = run(full, obs, grad=True, target=tgt, opts={"diagnostics": True})["stats"]
    estimators = {"pathwise": None,
                  "restart_every_window": {"tangent_windows": 1},
                  "frozen_noise": {"frozen_noise": True},
                  "frozen_noise_restart_every_window": {"frozen_noise": True,
                                                        "tangent_windows": 1},
                  "frozen_noise_graze_cut_0.1_restart_every_window": {
                      "frozen_noise": True, "tangent_windows": 1, "graze_cut": 0.1}}
    g = {k: {p: [] for p in params} for k in estimators}
    fd = {h: {p: [] for p in params} for h in rel_steps}
    for s in range(seeds):
        for h in rel_steps:
            for p in params:
                fd[h][p].append(central(full, obs, tgt, p, h, s))
        for k, o in estimators.items():
            v = pick(run(full, obs, grad=True, target=tgt, seed=s, opts=o))
            for p in params:
                g[k][p].append(v[p])
        print(f"[bench] adjoint seed {s}: " + "; ".join(
            f"{p}: central " + ", ".join(f"{fd[h][p][-1]:.4e}" for h in rel_steps) + ", "
            + ", ".join(f"{k} {g[k][p][-1]:.4e}" for k in estimators) for p in params),
            file=sys.stderr, flush=True)

    def stats(x, refs=None):
        x = np.asarray(x)
        out = {"mean": float(x.mean()), "std_error": float(x.std(ddof=1) / np.sqrt(len(x))),
               "per_seed": x.tolist()}
        for h, ref in (refs or {}).items():
            d = x - np.asarray(ref)
            out[f"bias_vs_central_{h:g}"] = {
                "mean": float(d.mean()), "std_error": float(d.std(ddof=1) / np.sqrt(len(d)))}
        return out

    checks = {}
    for p in params:
        refs = {h: fd[h][p] for h in rel_steps}
        checks[p] = {"central_difference": {f"rel_step_{h:g}": stats(fd[h][p])
                                            for h in rel_steps},
                     "estimators": {k: stats(g[k][p], refs) for k in estimators}}
    return {"workload": "config-1 field, 8 x-bands of 25 columns with own aL, aT; Dm and "
                        "porosity scale; 1e6 particles, 300 RW steps of 0.25 d, 12 "
                        "checkpoint windows; misfit at 32 observation points",
            "particles": n, "steps": total, "windows": len(full),
            "forward_ms": fwd["ms"], "adjoint_ms": A["ms"],
            "adjoint_over_forward": A["ms"] / fwd["ms"],
            "misfit": A["misfit"],
            "fd_check_4_steps": {p: {"adjoint": g_short[p],
                                     "central_difference": fd_short[p][-1]["central_difference"],
                                     "rel_step": fd_short[p][-1]["rel_step"],
                                     "ratio": g_short[p] / fd_short[p][-1]["central_difference"],
                                     "steps_tried": fd_short[p]}
                                 for p in params},
            "fd_check_300_steps": {"seeds": seeds, **checks},
            "sweep_diagnostics": diag}




def timed_run(P, rel, precision, steps, warmup, tuning_names, kde=True,
              adaptive_loops=0, graph=0, graph_max_list=0):
    """`warmup` intervals, then `steps` timed intervals from the release,
    each one rwpt_advance by 50 RW steps followed (kde) by one gpkde density
    pass of the active particles on the flow grid.  After RUN_INTERVALS
    intervals the run starts again from the release.  With adaptive_loops
    > 0 the gpkde optimisation is measured on the plume after the first
    interval before the timed region.  graph > 0 runs the walk as replays of
    captured graphs of that many launches (rwpt_set_graph) once at most
    graph_max_list particles are unfinished."""
    from rwpt import Engine
    E = Engine(P, precision=precision)
    tune = apply_tuning(E, tuning_names[precision], TUNED_NAME_1[precision])
    if graph:
        E.set_graph(graph, graph_max_list)
    nbins = (P.ncol, P.nrow, P.nlay)
    origin = (0.0, 0.0, float(P.bot.min()))

    def restart():
        E.load_particles(*rel)
        E.set_time(0.0)

    try:
        restart()
        for w in range(warmup):
            j = w % RUN_INTERVALS
            if j == 0 and w > 0:
                restart()
            E.advance((j + 1) * INTERVAL_STEPS * DT)
        adaptive = None
        if adaptive_loops > 0:
            # the plume after the first tracking interval, still inside the grid
            restart()
            E.advance(INTERVAL_STEPS * DT)
            adaptive = adaptive_kde(E, origin, nbins, adaptive_loops)
        restart()
        E.reset_kernel_time()
        total_steps = 0
        walk_s = kde_s = 0.0
        t0 = time.perf_counter()
        for k in range(steps):
            j = k % RUN_INTERVALS
            if j == 0 and k > 0:
                restart()
            ta = time.perf_counter()
            _, ns = E.advance((j + 1) * INTERVAL_STEPS * DT)
            tb = time.perf_counter()
            if kde:
                E.kde(origin, (1.0, 1.0, 1.0), nbins, mode=1)
            kde_s += time.perf_counter() - tb
            walk_s += tb - ta
            total_steps += ns
        el = time.perf_counter() - t0
        kms, launches = E.kernel_time(0)
        hms, _ = E.kernel_time(1)
        dms, _ = E.kernel_time(2)
        bw = E.kde_info()["bandwidth"]
        active = int((E.particles()["status"] == 1).sum())
    finally:
        E.close()
    return {"adaptive": adaptive, "tuning": tune,
            "steps": total_steps, "walk_s": walk_s, "kde_s": kde_s, "elapsed": el,
            "kernel_ms": kms, "launches": launches, "hist_ms": hms, "dens_ms": dms,
            "bandwidth": bw, "active": active,
            "rate": total_steps / el, "walk_rate": total_steps / walk_s}


def precision_errors(P, rel, precision, report_steps=(50, 100, 200, 400)):
    """Errors of the FP32-local walk against the plain FP64 walk on the
    same particles and draws (untimed): plume mean and variance of the
    active particles (moments summed in FP64) and the gpkde concentration
    at 64 observation cells after each of report_steps RW steps while at
    least 1 % of the particles are in the grid.  The same quantities
    between two FP64 walks with independent random streams give the
    Monte Carlo floor of the concentration differences: FP32 and FP64
    trajectories follow the same draws but part ways at cell crossings."""
    from rwpt import Engine
    nbins = (P.ncol, P.nrow, P.nlay)
    origin = (0.0, 0.0, float(P.bot.min()))
    obs = observation_cells(P)
    n = rel[0].size
    snaps = {}
    for key, prec, offset in (("fp64", 0, 0), ("fp32", precision, 0), ("fp64_indep", 0, n)):
        E = Engine(P, precision=prec)
        apply_tuning(E, TUNED_NAME[prec], TUNED_NAME_1[prec])
        try:
            E.load_particles(*rel, stream_offset=offset)
            res = []
            for target in report_steps:
                E.advance(target * DT)
                g = E.particles()
                on = g["status"] == 1
                gx, gy, gz = E.global_xyz()
                x = np.stack([gx[on], gy[on], gz[on]])
                dens, hist = E.kde(origin, (1.0, 1.0, 1.0), nbins, mode=1)
                res.append({"steps": target, "active": int(on.sum()),
                            "mean": x.mean(axis=1) if on.any() else None,
                            "var": x.var(axis=1) if on.any() else None,
                            "obs": dens[obs], "obs_counts": hist[obs].astype(float)})
        finally:
            E.close()
        snaps[key] = res
    errs = []
    for a, b, c in zip(snaps["fp64"], snaps["fp32"], snaps["fp64_indep"]):
        if a["active"] < 0.01 * n:
            continue  # the plume has left the grid: nothing left to compare
        e = {"steps": a["steps"], "active_fp64": a["active"], "active_fp32": b["active"],
             "mean_rel_err": float(np.max(np.abs(b["mean"] - a["mean"]) / np.abs(a["mean"]))),
             "variance_rel_err": float(np.max(np.abs(b["var"] - a["var"]) / a["var"])),
             "mean_rel_diff_indep": float(np.max(np.abs(c["mean"] - a["mean"]) /
                                                 np.abs(a["mean"]))),
             "variance_rel_diff_indep": float(np.max(np.abs(c["var"] - a["var"]) / a["var"])),
             "obs_particles_fp64": int(a["obs_counts"].sum())}
        if a["obs_counts"].sum() >= 100:  # else the plume has passed the observation cells
            e.update({
                "obs_conc_rel_l1": float(np.abs(b["obs"] - a["obs"]).sum() / a["obs"].sum()),
                "obs_conc_rel_l1_indep": float(np.abs(c["obs"] - a["obs"]).sum() /
                                               a["obs"].sum()),
                "obs_conc_rel_max": float(np.abs(b["obs"] - a["obs"]).max() / a["obs"].max()),
                "obs_count_rel_l1": float(np.abs(b["obs_counts"] - a["obs_counts"]).sum() /
                                          a["obs_counts"].sum()),
                "obs_count_rel_l1_indep": float(np.abs(c["obs_counts"] - a["obs_counts"]).sum()
                                                / a["obs_counts"].sum())})
        errs.append(e)
    worst = max((max(e["mean_rel_err"], e["variance_rel_err"]) for e in errs), default=None)
    gate = {"bound": ERROR_BOUND, "quantities": "plume mean and variance (concentration "
                                                "moments) of the FP32-local walk vs FP64",
            "worst": worst, "passed": worst is not None and worst <= ERROR_BOUND}
    return {"errors": errs, "error_gate": gate,
            "how": "FP32-local vs FP64 walk on the same particles and draws after 50, 100, "
                   "200, 400 RW steps (while >= 1 % of the particles are in the grid); "
                   "*_indep: FP64 vs FP64 with independent random streams (stream offset "
                   "1e7), the Monte Carlo floor of each difference; concentrations: one "
                   "gpkde density pass at 64 observation cells (8 x 8, middle layer), "
                   "relative L1 over the 64 cells while they hold >= 100 particles"}


def graph_run(P, rel, prec, steps, warmup, small=200_000, launches=8, max_list=200_000):
    """The walk with graph replays (rwpt_set_graph) against eager launches,
    walk only: the headline workload with the tail of each interval (at most
    max_list unfinished particles) replayed as graphs of `launches`
    launches, and the first `small` particles of it (a launch-bound run)
    replayed from the first launch."""
    out = {"launches_per_graph": launches}
    for key, sub, ml in (("config2_tail", None, max_list), ("small_full", small, 0)):
        r_rel = rel if sub is None else tuple(a[:sub] for a in rel)
        e = timed_run(P, r_rel, prec, steps, warmup, TUNED_NAME, kde=False)
        g = timed_run(P, r_rel, prec, steps, warmup, TUNED_NAME, kde=False, graph=launches,
                      graph_max_list=ml)
        assert g["steps"] == e["steps"]
        out[key] = {"particles": int(r_rel[0].size), "max_list": ml,
                    "eager_ms_per_step": e["elapsed"] * 1e3 / steps,
                    "graph_ms_per_step": g["elapsed"] * 1e3 / steps,
                    "speedup": e["elapsed"] / g["elapsed"]}
    return out


def usg_run(particles=1_000_000, intervals=10, dt=0.5):
    """The unstructured walk (rwpt_create_usg, FP64) on a smoothed quadtree
    of 60 x 60 base cells with a refined 10 x 10 block (2 x 2 sub-cells), 10
    layers, 39,000 cells, stream-function flow, 1e6 particles over the grid,
    `intervals` intervals of 50 RW steps; beside it the same base grid
    unrefined through the unstructured engine and through the structured
    one (the same cells and flows)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from rwpt import Engine, problems
    from usg_synth import Quadtree
    from usg_scenarios import release
    from test_oracle_usg_vs_reference import _structured_twin
    out = {"workload": "DISV quadtree 60x60 base cells, refined 10x10 block, 10 layers, "
                       "%d particles, %d intervals of 50 RW steps, FP64" % (particles,
                                                                             intervals)}
    for key, refine in (("refined", {(i, j) for i in range(25, 35) for j in range(20, 30)}),
                        ("unrefined", set())):
        Q = Quadtree(60, 60, refine, nlay=10)
        U = problems.UsgProblem(ncell=Q.ncell, nlay=Q.nlay, ia=Q.ia, ja=Q.ja,
                                flowja=Q.flows(seed=3), porosity=np.full(Q.ncell, 0.3),
                                aL=np.array([0.3]), aT=np.array([0.03]),
                                dmeff=np.array([1e-4]), ts_params=(dt, 0.0),
                                **Q.cell_arrays())
        rel = release(U, particles, seed=5)
        probs = [("usg", U)] + ([("structured", _structured_twin(Q, U))] if not refine else [])
        for name, prob in probs:
            E = Engine(prob)
            try:
                E.load_particles(*rel)
                E.advance(INTERVAL_STEPS * dt)          # warm-up interval
                E.reset_kernel_time()
                steps = 0
                t0 = time.perf_counter()
                for k in range(2, intervals + 2):
                    steps += E.advance(k * INTERVAL_STEPS * dt)[1]
                el = time.perf_counter() - t0
                kms, launches = E.kernel_time(0)
            finally:
                E.close()
            out[f"{key}_{name}"] = {"value": steps / el, "unit": "particle-steps/s",
                                    "cells": int(Q.ncell), "steps": int(steps),
                                    "ms_per_interval": el * 1e3 / intervals,
                                    "walk_kernel_avg_launch_ms": kms / max(launches, 1)}
    return out


def config1_run(steps=10):
    """The config-1 workload (1e6 particles, 10 intervals): walk rates of
    the FP32-local and the FP64 walk."""
    P, rel = workload(1_000_000)
    out = {}
    for prec in (1, 0):
        r = timed_run(P, rel, prec, steps, 1, TUNED_NAME_1, kde=False)
        out["fp32" if prec else "fp64"] = {
            "value": r["rate"], "particle_steps": r["steps"], "elapsed_s": r["elapsed"],
            "walk_kernel_avg_launch_ms": r["kernel_ms"] / max(r["launches"], 1),
            "launches": r["launches"],
            "roofline_frac": BYTES_PER_STEP[prec] * r["steps"] / (r["kernel_ms"] / 1e3)
            / 1e9 / HBM_PEAK_GBS}
    out["speedup_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["workload"] = "config-1 field, 1e6 particles in the inlet block, 10 intervals of 50 RW steps"
    return out


def read_traffic(prec):
    tpath = os.path.join(ROOT, "profiles", TRAFFIC_FILE[prec])
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":")[1]), TRAFFIC_FILE[prec]
    return None, "not profiled for this precision"


def parse_args(argv=None):
    """The command line; exits with status 2 on a request this build cannot
    meet (more than one GPU, no timed step)."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=RUN_INTERVALS,
                    help="timed tracking intervals of 50 RW steps (20 = the 1000-step "
                         "run; more restart from the release)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=HEADLINE_PARTICLES)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000,
                    help="particles of the oracle's CPU-baseline interval")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", type=int, default=1, choices=(0, 1, 2),
                    help="walk arithmetic: 0 plain FP64, 1 FP32-local with the FP64 "
                         "walk's draws (default), 2 FP32-local with 4-draw normals")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the plain FP64 comparison run and the error pass")
    ap.add_argument("--quick", action="store_true",
                    help="headline only: no config1, ensemble, adjoint or adaptive "
                         "gpkde legs")
    ap.add_argument("--ensemble-intervals", type=int, default=1,
                    help="tracking intervals of the 256-member ensemble timed (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint (calibration) workload")
    ap.add_argument("--kde-loops", type=int, default=10,
                    help="optimisation loops of the adaptive gpkde (0: skip)")
    ap.add_argument("--adjoint-seeds", type=int, default=8,
                    help="base seeds of the adjoint leg's central-difference check (>= 2)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint leg (its JSON object on one line)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device 3 times (the known-size "
                         "workload that calibrates rocprofv3's FETCH_SIZE and WRITE_SIZE)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every GPU box here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)
    if args.adjoint_seeds < 2:
        print("--adjoint-seeds must be >= 2", file=sys.stderr)
        sys.exit(2)
    if args.quick:
        args.ensemble_intervals, args.no_adjoint, args.kde_loops = 0, True, 0
    return args


def copy_probe(gib=1, reps=3):
    """A plain device-to-device copy of `gib` GiB, `reps` times (HIP's own
    copy kernel): the known-size workload that scales the FETCH_SIZE and
    WRITE_SIZE counters of a profile to HBM bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(gib << 30)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, n) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(reps):
            if hip.hipMemcpy(b, a, n, 3) != 0:   # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(json.dumps({"copy_probe": f"{reps} x {gib} GiB device-to-device"}), flush=True)


def main(argv=None):
    args = parse_args(argv)
    if args.copy_probe:
        copy_probe()
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_run(seeds=args.adjoint_seeds)), flush=True)
        return
    P, rel = workload(args.particles)
    prec = args.precision
    print(f"[bench] headline: {args.particles} particles, precision {prec}, "
          f"{args.warmup} warm-up + {args.steps} timed intervals", file=sys.stderr, flush=True)
    r = timed_run(P, rel, prec, args.steps, args.warmup, TUNED_NAME,
                  adaptive_loops=args.kde_loops)
    r0 = errors = None
    if prec != 0 and not args.no_fp64:
        print("[bench] FP64 comparison run", file=sys.stderr, flush=True)
        r0 = timed_run(P, rel, 0, args.steps, args.warmup, TUNED_NAME)
        print("[bench] FP32 vs FP64 error pass", file=sys.stderr, flush=True)
        errors = precision_errors(P, rel, prec)
    value, value_from = r["rate"], PRECISION_NAME[prec]
    if errors is not None and not errors["error_gate"]["passed"]:
        value, value_from = r0["rate"], PRECISION_NAME[0] + " (FP32 error gate failed)"
    # roofline of the walk kernel over the timed region's launches
    avg_launch_s = (r["kernel_ms"] / 1e3) / max(r["launches"], 1)
    bytes_per_launch = BYTES_PER_STEP[prec] * r["steps"] / max(r["launches"], 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic, traffic_src = read_traffic(prec)
    nbins = (P.ncol, P.nrow, P.nlay)
    kde_cells = int(np.prod(nbins))
    out = {
        "metric": "particle-steps/s + KDE cells/s, 1 GPU: FP32/batched vs plain "
                  "FP64 at <=0.5% err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["elapsed"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32",
        "data": "synthetic (seeded lognormal field, curl-of-potential fluxes)",
        "config": {
            "workload": "config 2: heterogeneous 200x200x20 lognormal field, %d particles "
                        "in the 10x50x20-cell inlet block, FIXED dt 0.5 d, one step = one "
                        "tracking interval of 50 RW steps + one gpkde density pass, 20 "
                        "intervals = 1000 RW steps" % args.particles,
            "grid": [200, 200, 20], "particles": args.particles,
            "rw_steps_per_interval": INTERVAL_STEPS, "dt_days": DT,
            "precision": PRECISION_NAME[prec], "advection": "EULERIAN",
        },
        "value_from": value_from,
        "particle_steps": r["steps"],
        "walk_only_rate": r["walk_rate"],
        "active_after": r["active"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "walk_kernel",
                     "launches": r["launches"],
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_bytes_per_step": BYTES_PER_STEP[prec],
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "tuning": r["tuning"],
        "kde": {"unit": "KDE cells/s", "grid": list(nbins),
                "reconstructions": args.steps,
                "cells_per_s_wall": kde_cells * args.steps / r["kde_s"],
                "cells_per_s_kernels": kde_cells * args.steps /
                ((r["hist_ms"] + r["dens_ms"]) / 1e3),
                "histogram_ms": r["hist_ms"] / args.steps,
                "density_ms": r["dens_ms"] / args.steps,
                "wall_ms": r["kde_s"] * 1e3 / args.steps,
                "last_bandwidth": r["bandwidth"]},
    }
    if r["adaptive"] is not None:
        out["kde_adaptive"] = r["adaptive"]
    if r0 is not None:
        out["fp64"] = {"value": r0["rate"], "walk_only_rate": r0["walk_rate"],
                       "ms_per_step": r0["elapsed"] * 1e3 / args.steps,
                       "particle_steps": r0["steps"],
                       "walk_kernel_avg_launch_ms": r0["kernel_ms"] / max(r0["launches"], 1),
                       "launches": r0["launches"],
                       "roofline_frac": BYTES_PER_STEP[0] * r0["steps"] /
                       (r0["kernel_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS}
        out["speedup_vs_fp64"] = r["rate"] / r0["rate"]
        out["error_vs_fp64"] = errors
    if not args.quick:
        print("[bench] graph replay leg", file=sys.stderr, flush=True)
        out["graph"] = graph_run(P, rel, prec, args.steps, args.warmup)
        print("[bench] config1 leg", file=sys.stderr, flush=True)
        out["config1"] = config1_run()
        print("[bench] unstructured-grid leg", file=sys.stderr, flush=True)
        out["usg"] = usg_run()
    if args.ensemble_intervals > 0:
        print("[bench] ensemble leg", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_run(args.ensemble_intervals, 1)
    if not args.no_adjoint:
        print("[bench] adjoint leg", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_run(seeds=args.adjoint_seeds)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(P, rel, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
