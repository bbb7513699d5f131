"""The dual-number oracle (oracle/rwpt_oracle_dual.cpp: rwpt_oracle.c with
every double a dual number), the FP64 reference gradient of the adjoint
tests: its values equal the plain oracle's bit for bit (the branches are
the oracle's), and its tangents of every particle's final position, along
aL, aT, Dm in every cell and along a porosity scale, equal central
differences of the plain oracle with common random numbers for the
particles whose branches the +-h runs share (cells, statuses and sub-step
counts), to the differences' own truncation and rounding error."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from oracle import oracle_dual as D
from rwpt import problems

WINDOWS = (1.0, 2.0, 3.0)


def _problem():
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.ts_params = (0.5, 0.0)
    P.time_step_kind = 4
    zones = problems.column_zones(P, 3)
    Q = problems.zoned_dispersion(P, zones, [0.3, 0.5, 0.4], [0.04, 0.05, 0.03], 1e-3)
    rel = problems.release_in_block(P, 400, (2, 6), (6, 15), (1, 6), seed=5)
    return Q, rel


def _tangents(Q):
    t = {k: np.zeros((Q.ncell, D.ND)) for k in ("aL", "aT", "dmeff", "porosity")}
    t["aL"][:, 0] = 1.0
    t["aT"][:, 1] = 1.0
    t["dmeff"][:, 2] = 1.0
    t["porosity"][:, 3] = Q.porosity
    return t


def _run_plain(Q, rel):
    st = O.new_state(*rel)
    OP = O.OracleProblem(Q)
    t0 = 0.0
    for t1 in WINDOWS:
        O.advance(OP, st, t0, t1, nthreads=1)
        t0 = t1
    return st


@pytest.fixture(scope="module")
def duals():
    Q, rel = _problem()
    DP = D.DualProblem(Q, _tangents(Q))
    st = D.new_state(*rel)
    t0 = 0.0
    for t1 in WINDOWS:
        assert D.advance(DP, st, t0, t1) == 0
        t0 = t1
    return Q, rel, st


def test_values_equal_the_oracle(duals):
    Q, rel, st = duals
    ref = _run_plain(Q, rel)
    for k in ("cell", "status", "substeps", "rng"):
        assert np.array_equal(st[k], ref[k]), k
    for k in ("lx", "ly", "lz", "t"):
        assert np.array_equal(st[k][:, 0], ref[k], equal_nan=True), k
    assert len(np.unique(st["status"])) >= 2


@pytest.mark.parametrize("k,name,hrel", [(0, "aL", 1e-7), (1, "aT", 1e-7),
                                         (2, "dmeff", 1e-6), (3, "porosity", 1e-8)])
def test_tangents_equal_central_differences(duals, k, name, hrel):
    Q, rel, st = duals
    runs = []
    for sgn in (1.0, -1.0):
        Qx = dataclasses.replace(Q)
        if name == "porosity":
            Qx.porosity = Q.porosity * (1.0 + sgn * hrel)
            h = hrel
        else:
            base = getattr(Q, name)
            h = hrel * float(np.mean(base))
            setattr(Qx, name, base + sgn * h)
        runs.append(_run_plain(Qx, rel))
    a, b = runs
    same = ((a["cell"] == b["cell"]) & (a["status"] == b["status"]) &
            (a["substeps"] == b["substeps"]))
    assert same.mean() > 0.9
    for f in ("lx", "ly", "lz", "t"):
        fd = (a[f] - b[f]) / (2.0 * h)
        tg = st[f][:, 1 + k]
        err = np.abs(fd - tg)[same] / (np.abs(tg[same]) + 1e-3)
        assert np.median(err) < 1e-6 and np.quantile(err, 0.9) < 1e-5, f
        assert err.max() < 1e-2, f
    # the directions move the particles
    assert np.abs(st["lx"][:, 1 + k]).max() > 1.0
