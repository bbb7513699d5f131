"""Benchmark of the random-walk hot path on one MI355X.

Workload (one "step" = one tracking interval = one rwpt_advance call):
the heterogeneous 200x200x20 lognormal field (1 m cells, divergence-free
face flows, porosity U[0.20,0.30], aL=0.5 m, aT=0.05 m, Dm=1e-4 m2/d),
1e6 particles released uniformly in the 10x50x20-cell inlet block, FIXED
dt = 0.5 d, timeseries every 50 steps -- so one bench step advances every
particle by 50 RW time steps (25 d) and --steps 10 is the whole 500-step
run.  Particle state is resident in HBM before the timed region.

Each interval ends with the gpkde density reconstruction of the active
particles on the 200x200x20 flow grid (timed apart, reported as KDE cells/s;
the wall time includes the host copies of the C ABI).

Prints one JSON line: particle-steps/s (accepted RW sub-steps, the unit the
reference's TrackSubCell loop iterates, over the walk's wall time), the roofline of the advance kernel
(algorithmic bytes = 96 B of SoA state per particle-step, HBM bound, 8 TB/s)
and the CPU baseline (the oracle, OpenMP, on this host's cores, on a bounded
sample of the same workload).
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

BYTES_PER_STEP_F64 = 96   # lx,ly,lz f64 + cell,status i32 + 2 x u64, read+write
HBM_PEAK_GBS = 8000.0
INTERVAL_STEPS = 50
DT = 0.5


def workload(n_particles):
    from rwpt import problems
    P = problems.config1()
    rel = problems.release_in_block(P, n_particles, (1, 10), (76, 125), (1, 20))
    return P, rel


def cpu_baseline(P, rel, n_sample, seconds_budget=20.0):
    """The oracle (plain C, OpenMP over particles) on this host's cores: the
    first tracking interval of the same workload on n_sample particles."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    cell, lx, ly, lz = (a[:n_sample] for a in rel)
    st = O.new_state(cell, lx, ly, lz)
    OP = O.OracleProblem(P)
    t0 = time.perf_counter()
    O.advance(OP, st, 0.0, INTERVAL_STEPS * DT, nthreads=threads)
    el = time.perf_counter() - t0
    steps = int(st["substeps"].sum())
    return {"value": steps / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{n_sample} particles x {INTERVAL_STEPS} RW steps "
                      f"({steps} particle-steps) of the same workload, "
                      f"{el:.1f} s on {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every GPU box here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.steps > 10:
        print("--steps must be 1..10 (the workload has 10 tracking intervals)",
              file=sys.stderr)
        sys.exit(2)

    from rwpt import Engine
    P, rel = workload(args.particles)
    E = Engine(P)
    cell, lx, ly, lz = rel

    # warm-up intervals from the initial state, then restart the same run
    E.load_particles(cell, lx, ly, lz)
    E.set_time(0.0)
    for w in range(args.warmup):
        E.advance((w + 1) * INTERVAL_STEPS * DT)
    E.load_particles(cell, lx, ly, lz)
    E.set_time(0.0)
    E.reset_kernel_time()

    # each interval ends with the gpkde reconstruction of the active plume on
    # the flow grid (1 m bins, Silverman bandwidth, reflected kernel)
    nbins = (P.ncol, P.nrow, P.nlay)
    kde_cells = int(np.prod(nbins))
    total_steps = 0
    walk_s = kde_s = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        ta = time.perf_counter()
        _, ns = E.advance((k + 1) * INTERVAL_STEPS * DT)
        tb = time.perf_counter()
        E.kde((0.0, 0.0, float(P.bot.min())), (1.0, 1.0, 1.0), nbins, mode=1)
        kde_s += time.perf_counter() - tb
        walk_s += tb - ta
        total_steps += ns
    el = time.perf_counter() - t0
    kms, launches = E.kernel_time(0)
    hms, _ = E.kernel_time(1)
    dms, _ = E.kernel_time(2)
    bw = E.kde_info()["bandwidth"]
    state = E.particles()
    E.close()

    value = total_steps / walk_s
    avg_launch_s = (kms / 1e3) / max(launches, 1)
    bytes_per_launch = BYTES_PER_STEP_F64 * total_steps / max(launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "walk_kernel_config1_traffic.txt")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("hbm_bytes_per_launch:"):
                traffic = float(line.split(":")[1])
    out = {
        "metric": "particle-steps/s + KDE cells/s, 1 GPU: FP32/batched vs plain "
                  "FP64 at <=0.5% err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal field, curl-of-potential fluxes)",
        "config": {
            "workload": "heterogeneous 200x200x20 lognormal field, 1e6 particles, "
                        "FIXED dt 0.5 d, 50 RW steps per tracking interval",
            "grid": [200, 200, 20], "particles": args.particles,
            "rw_steps_per_interval": INTERVAL_STEPS, "dt_days": DT,
            "precision": "plain FP64", "advection": "EULERIAN",
        },
        "particle_steps": total_steps,
        "active_after": int((state["status"] == 1).sum()),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": "walk_kernel",
                     "launches": launches,
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "kde": {"unit": "KDE cells/s", "grid": list(nbins),
                "reconstructions": args.steps,
                "cells_per_s_wall": kde_cells * args.steps / kde_s,
                "cells_per_s_kernels": kde_cells * args.steps / ((hms + dms) / 1e3),
                "histogram_ms": hms / args.steps, "density_ms": dms / args.steps,
                "wall_ms": kde_s * 1e3 / args.steps, "last_bandwidth": bw},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, rel, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
