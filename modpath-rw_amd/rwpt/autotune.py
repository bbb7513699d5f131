"""Launch-shape autotuning of the walk (rwpt_autotune) for named workloads,
kept in tuning.json beside this file and read back by tuned().

    python -m rwpt.autotune [--out PATH] [--reps N] [workload ...]

Each workload is built, warmed up for one tracking interval, and tuned with
trial advances over the next interval (rwpt_autotune restores the state
after every trial).  The walk's results do not depend on the shape, only
its speed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_FILE = os.path.join(HERE, "tuning.json")

INTERVAL_STEPS = 50
DT = 0.5


def _config1(precision, n=1_000_000):
    from . import problems
    P = problems.config1()
    rel = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    return P, rel, precision, INTERVAL_STEPS


def _ensemble(precision):
    from . import problems
    E, counts = problems.config3()
    rel = problems.release_ensemble(E, counts, (1, 5), (38, 63), (1, 10))
    return E, rel, precision, 10


WORKLOADS = {
    "config1_fp64": lambda: _config1(0),
    "config1_fp32": lambda: _config1(1),
    "config2_fp64": lambda: _config1(0, 10_000_000),
    "config2_fp32": lambda: _config1(1, 10_000_000),
    "ensemble256_fp32": lambda: _ensemble(1),
}


def tune(name, reps=3):
    """Best shape of `reps` autotuner passes (the fastest trial wins)."""
    from .engine import Engine
    P, rel, precision, probe_steps = WORKLOADS[name]()
    E = Engine(P, precision=precision)
    try:
        E.load_particles(*rel)
        E.advance(INTERVAL_STEPS * DT)
        best = None
        for _ in range(reps):
            r = E.autotune(INTERVAL_STEPS * DT + probe_steps * DT)
            if best is None or r["ms"] < best["ms"]:
                best = r
    finally:
        E.close()
    return {k: best[k] for k in ("budget", "block", "sort", "min_waves", "ms")}


def tuned(name):
    """The stored shape of a workload, or None."""
    if not os.path.exists(TUNING_FILE):
        return None
    return json.load(open(TUNING_FILE)).get(name)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--out", default=TUNING_FILE)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    table = json.load(open(a.out)) if os.path.exists(a.out) else {}
    for name in a.workloads:
        table[name] = tune(name, a.reps)
        print(name, table[name], flush=True)
        with open(a.out, "w") as fh:
            json.dump(table, fh, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
