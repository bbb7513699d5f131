"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

EULERIAN runs must agree bit for bit: every operation on that path is an
IEEE-754 add, multiply, divide or square root in the reference's operand
order, built without fused multiply-add on either side.  EXPONENTIAL runs
call exp/pow, whose last bit differs between the device and host math
libraries, so they are compared with a tolerance on positions (1e-10
relative) and on the fraction of particles whose path is unchanged.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from rwpt import Engine, problems
from scenarios import SCEN, release as _release

pytestmark = pytest.mark.gpu

FIELDS = ("cell", "status", "lx", "ly", "lz", "t", "substeps", "rng")


@pytest.mark.parametrize("name", sorted(SCEN))
def test_cell_tables_bitwise(name):
    P = SCEN[name][0]
    E = Engine(P)
    try:
        got = E.cell_tables()
    finally:
        E.close()
    want = O.cell_tables(P)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("name", sorted(SCEN))
def test_advance_bitwise_eulerian(name):
    P, where, dt_out, nout = SCEN[name]
    n = 1500
    cell, lx, ly, lz = _release(P, where, n)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        st = O.new_state(cell, lx, ly, lz)
        OP = O.OracleProblem(P)
        t = 0.0
        for k in range(1, nout + 1):
            t1 = k * dt_out
            act, ns = E.advance(t1)
            before = st["substeps"].sum()
            assert O.advance(OP, st, t, t1) == 0
            assert ns == st["substeps"].sum() - before
            assert act == int((st["status"] == 1).sum())
            t = t1
        g = E.particles()
    finally:
        E.close()
    for f in FIELDS:
        nan = f in ("lx", "ly", "lz", "t")
        assert np.array_equal(g[f], st[f], equal_nan=nan), \
            (f, int((g[f] != st[f]).sum()))
    # the walk did something
    assert st["substeps"].sum() > n * nout


def test_pending_release_and_unreleased():
    P, where, _, _ = SCEN["hetero3d"]
    n = 600
    cell, lx, ly, lz = _release(P, where, n)
    t0 = np.random.default_rng(1).uniform(-1.0, 6.0, n)
    status = np.zeros(n, np.int32)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz, t=t0, status=status)
        st = O.new_state(cell, lx, ly, lz)
        st["t"] = t0.copy()
        st["status"] = status.copy()
        OP = O.OracleProblem(P)
        t = 0.0
        for k in range(1, 5):
            E.advance(k * 2.0)
            O.advance(OP, st, t, k * 2.0)
            t = k * 2.0
        g = E.particles()
    finally:
        E.close()
    for f in FIELDS:
        assert np.array_equal(g[f], st[f]), f
    assert (g["status"] == 8).sum() == (t0 < 0).sum()


def test_rng_streams_match_jumps():
    P = SCEN["cfg0"][0]
    n = 300
    E = Engine(P)
    try:
        c = np.full(n, 2410, np.int32)
        E.load_particles(c, np.full(n, .5), np.full(n, .5), np.full(n, .5))
        a = E.particles()["rng"]
        E.load_particles(c, np.full(n, .5), np.full(n, .5), np.full(n, .5),
                         stream_offset=100000)
        b = E.particles()["rng"]
    finally:
        E.close()
    assert np.array_equal(a, O.rng_streams(n))
    assert np.array_equal(b[:20], O.rng_streams(10, first=100000))


@pytest.mark.parametrize("name", ["cfg0", "hetero3d", "hetero_ibound_rebound_R"])
def test_advance_exponential_close(name):
    P, where, dt_out, nout = SCEN[name]
    P = problems.Problem(**{**P.__dict__, "advection_kind": 1})
    n = 1500
    cell, lx, ly, lz = _release(P, where, n)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        st = O.new_state(cell, lx, ly, lz)
        OP = O.OracleProblem(P)
        t = 0.0
        for k in range(1, nout + 1):
            E.advance(k * dt_out)
            O.advance(OP, st, t, k * dt_out)
            t = k * dt_out
        g = E.particles()
    finally:
        E.close()
    same = (g["cell"] == st["cell"]) & (g["status"] == st["status"]) & \
        np.all(g["rng"].reshape(-1, 2) == st["rng"].reshape(-1, 2), axis=1)
    assert same.mean() >= 0.99
    for f in ("lx", "ly", "lz", "t"):
        d = np.abs(g[f][same] - st[f][same]) / np.maximum(1.0, np.abs(st[f][same]))
        assert d.max() <= 1e-10, f


def test_full_size_properties():
    """Config-1 field at full size: sizes the oracle cannot follow, checked
    through invariants -- determinism, state domains, time at t_stop."""
    P = problems.config1()
    n = 1_000_000
    cell, lx, ly, lz = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        act, ns = E.advance(25.0)
        a = E.particles()
        E.load_particles(cell, lx, ly, lz)
        act2, ns2 = E.advance(25.0)
        b = E.particles()
    finally:
        E.close()
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    assert act == act2 and ns == ns2 and ns > 10 * n
    on = a["status"] == 1
    assert act == on.sum() and on.mean() > 0.5
    assert np.all(a["t"][on] == 25.0)
    for f in ("lx", "ly", "lz"):
        assert np.all((a[f] >= 0.0) & (a[f] <= 1.0))
    assert np.all((a["cell"] >= 1) & (a["cell"] <= P.ncell))
    assert set(np.unique(a["status"])) <= {1, 2, 9}
    assert (a["status"] == 9).mean() < 1e-4


@pytest.mark.parametrize("budget,block,sort,minw", [(1, 64, 1, 1), (7, 128, 0, 1),
                                                    (64, 256, 1, 3), (100000, 256, 1, 4)])
def test_walk_launch_shape_does_not_change_results(budget, block, sort, minw):
    """The walk pauses each particle after `budget` iterations and resumes
    it in the next launch from its saved loop state; any budget, block size
    and list order must reproduce the oracle bit for bit."""
    P, where, dt_out, nout = SCEN["hetero_ibound_rebound_R"]
    n = 1500
    cell, lx, ly, lz = _release(P, where, n)
    E = Engine(P)
    try:
        E.set_tuning(budget, block, sort, minw)
        E.load_particles(cell, lx, ly, lz)
        st = O.new_state(cell, lx, ly, lz)
        OP = O.OracleProblem(P)
        t = 0.0
        for k in range(1, 4):
            E.advance(k * dt_out)
            if budget == 1:
                assert E.last_launches() > 10
            O.advance(OP, st, t, k * dt_out)
            t = k * dt_out
        g = E.particles()
    finally:
        E.close()
    for f in FIELDS:
        assert np.array_equal(g[f], st[f]), f


def test_steady_state_follows_the_loaded_step():
    """rwpt_set_steady_state: an engine created for steady flow and switched
    to a transient step walks as the oracle does with steady_state 0 (cells
    without an exit face keep their particles when stop-at-weak-sinks is on
    and the sink flow is zero, TrackCell.f90:464-497), and differently from
    the steady walk."""
    P, where, dt_out, nout = SCEN["transient_weak_sinks"]
    assert P.steady_state == 0
    n = 1500
    cell, lx, ly, lz = _release(P, where, n)
    E = Engine(dataclasses.replace(P, steady_state=1))
    try:
        E.set_steady_state(0)
        E.load_particles(cell, lx, ly, lz)
        for k in range(1, nout + 1):
            E.advance(k * dt_out)
        g = E.particles()
    finally:
        E.close()
    runs = {}
    for steady in (0, 1):
        st = O.new_state(cell, lx, ly, lz)
        OP = O.OracleProblem(dataclasses.replace(P, steady_state=steady))
        t = 0.0
        for k in range(1, nout + 1):
            assert O.advance(OP, st, t, k * dt_out) == 0
            t = k * dt_out
        runs[steady] = st
    for f in FIELDS:
        assert np.array_equal(g[f], runs[0][f], equal_nan=f in ("lx", "ly", "lz", "t")), f
    assert (runs[1]["status"] == 5).sum() > 0 and (runs[0]["status"] == 5).sum() == 0

