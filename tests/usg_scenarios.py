"""Named unstructured-grid problems (MODFLOW 6 DISV smoothed quadtrees) for
the oracle-vs-reference tests and the GPU parity tests.

Each scenario is a Quadtree (tests/usg_synth.py) with flows and options;
build(name) returns (Quadtree, rwpt.problems.UsgProblem).  The options
cover what the unstructured path adds to the structured one: refined cells
(sub-cells, internal faces, neighbours two cells wide), cells that meet a
bigger cell, several layers, boundary flows and weak sinks and sources on
split faces, reflecting boundaries, porosity that varies (the reference's
corner porosity is then NaN in some sub-cells of refined cells, see
oracle/rwpt_oracle_usg.c), backward tracking, the axisymmetric model,
exponential advection and the time-step kinds.
"""
import numpy as np

from rwpt.problems import UsgProblem
from usg_synth import Quadtree

GRIDS = {
    "quad1": dict(nrow=4, ncol=5, refine={(1, 2), (2, 1)}, nlay=1),
    "quad3": dict(nrow=5, ncol=6, refine={(0, 0), (0, 1), (1, 1), (2, 3), (4, 5)}, nlay=3),
    "block": dict(nrow=6, ncol=6, refine={(2, 2), (2, 3), (3, 2), (3, 3)}, nlay=2),
    "checker": dict(nrow=6, ncol=7,
                    refine={(i, j) for i in range(6) for j in range(7) if (i + j) % 3 == 0},
                    nlay=2),
    "plain": dict(nrow=5, ncol=6, refine=set(), nlay=2),
}

SCENARIOS = {
    # name: (grid, flow seed, option overrides)
    "quad1_2d": ("quad1", 0, dict(dim_mask=(1, 1, 0))),
    "quad1": ("quad1", 0, {}),
    "quad3": ("quad3", 1, {}),
    "block": ("block", 2, {}),
    "checker": ("checker", 3, {}),
    "quad3_reflect_porosity": ("quad3", 1, dict(default_icbound=1, porosity="random")),
    "block_sinks": ("block", 2, dict(bflow="random", sinks="random", stop_weak_sinks=1,
                                      steady_state=0)),
    "checker_backward_axi": ("checker", 3, dict(backward=1, disp_model=1, sources="random",
                                                stop_weak_sources=1, time_step_kind=3,
                                                ts_params=(0.2, 0.1))),
    "quad3_exponential": ("quad3", 1, dict(advection_kind=1, time_step_kind=1,
                                           ts_params=(0.1, 0.0))),
    "block_zones": ("block", 2, dict(zones="random", stop_zone=2, storage="random")),
}


def build(name, seed=11):
    grid, fseed, opt = SCENARIOS[name]
    Q = Quadtree(**GRIDS[grid])
    rng = np.random.default_rng(seed)
    A = Q.cell_arrays()
    n = Q.ncell
    kw = dict(opt)
    por = kw.pop("porosity", None)
    porosity = rng.uniform(0.2, 0.35, n) if por == "random" else np.full(n, 0.3)
    U = UsgProblem(ncell=n, nlay=Q.nlay, ia=Q.ia, ja=Q.ja, flowja=Q.flows(seed=fseed),
                   porosity=porosity, aL=np.array([0.3]), aT=np.array([0.03]),
                   dmeff=np.array([1e-4]), ts_params=(0.1, 0.0), **A)
    if kw.pop("bflow", None) == "random":
        bf = np.zeros(6 * n)
        idx = rng.choice(6 * n, max(4, n // 4), replace=False)
        bf[idx] = rng.normal(0.0, 0.3, idx.size)
        U.bflow = bf
    if kw.pop("sinks", None) == "random":
        U.sink = np.where(rng.random(n) < 0.15, -rng.uniform(0.01, 0.2, n), 0.0)
    if kw.pop("sources", None) == "random":
        U.source = np.where(rng.random(n) < 0.15, rng.uniform(0.01, 0.2, n), 0.0)
    if kw.pop("storage", None) == "random":
        U.storage = rng.normal(0.0, 0.02, n)
    if kw.pop("zones", None) == "random":
        U.zone = rng.integers(1, 4, n).astype(np.int32)
    if kw.get("disp_model") == 1:
        U.aL = rng.uniform(0.1, 0.4, n)
        U.aT = 0.1 * U.aL
        U.aLV = 0.5 * U.aL
        U.aTV = 0.5 * U.aT
        U.dmeff = np.full(n, 1e-4)
    for k, v in kw.items():
        setattr(U, k, v)
    return Q, U


def release(U, n, seed=3):
    """n particles at uniform random locations of uniform random cells."""
    r = np.random.default_rng(seed)
    cell = r.integers(1, U.ncell + 1, n).astype(np.int32)
    return cell, r.random(n), r.random(n), r.random(n)
