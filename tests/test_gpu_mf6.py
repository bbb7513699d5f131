"""MODFLOW 6 DIS input through the engine on the GPU.

The engine is built from rwpt.mf6's reading of a small MODFLOW 6 DIS
simulation's grid, head and budget files (the scenario of
tests/test_mf6.py).  Its cell records must equal, bit for bit, the
reference's MODFLOW 6 assembly of the same files (face pore velocities,
DX DY DZ, source and sink flows; oracle/_ref/mf6_ref) and the C oracle's
records, and a walk over them must equal the oracle's walk bit for bit
(EULERIAN: IEEE add, multiply, divide and square root in the reference's
operand order on both sides).  A driver over a two-step MODFLOW 6 run
loads each step's flow before the step's first interval."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import refrun
from rwpt import Engine, mf6, problems

from test_mf6 import (HDRY, IFACE, REMOVED, engine_problem, oracle_problem, scenario,
                      two_step_scenario)

pytestmark = pytest.mark.gpu

FIELDS = ("cell", "status", "lx", "ly", "lz", "t", "substeps", "rng")
# the scenario as written, and with cells removed by IDOMAIN 0
GRIDS = pytest.mark.parametrize("removed", [(), tuple(REMOVED)], ids=["full", "idomain"])


def _read(tmp, removed=()):
    paths, P0, ibound = scenario(tmp, removed=removed)
    por = np.full(P0.ncell, 0.25)
    por[::5] = 0.3
    ret = np.ones(P0.ncell)
    ret[::7] = 1.5
    g = mf6.read_grb(paths[0])
    ts = mf6.load_time_step(g, mf6.read_heads(paths[1]), mf6.read_budget(paths[2]), 1, 1,
                            ibound, HDRY, dict(IFACE))
    P, extras = mf6.problem(g, ts, por, ret, aL=P0.aL, aT=P0.aT, dmeff=P0.dmeff,
                            ts_params=(0.1, 0.0))
    Q, _ = engine_problem(P0, g, ts, por, ret)
    for k in ("frf", "fff", "flf", "bflow", "ibound", "top", "bot", "delr", "delc"):
        assert np.array_equal(getattr(P, k), getattr(Q, k)), k
    return paths, P, extras, ibound, por, ret


def _engine(P, extras):
    E = Engine(P)
    E.set_cell_extras(**extras)
    return E


@GRIDS
def test_engine_cells_match_reference_mf6(tmp_path, removed):
    paths, P, extras, ibound, por, ret = _read(str(tmp_path), removed)
    print("mf6 cells", P.ncol, P.nrow, P.nlay, flush=True)
    E = _engine(P, extras)
    try:
        got = E.cell_tables()
    finally:
        E.close()
    want = O.cell_tables(oracle_problem(P, extras))
    assert np.array_equal(got, want), np.abs(got - want).max()
    if not refrun.mf6_available():
        pytest.skip("oracle/_ref/mf6_ref is not built")
    ref = refrun.mf6(*paths, ibound, por, ret, hdry=HDRY, default_iface=IFACE)
    act = P.ibound != 0
    np.testing.assert_array_equal(got[act, 0:6], ref["velocity"][act])
    np.testing.assert_array_equal(got[act, 6:9], ref["size"][act])
    np.testing.assert_array_equal(got[act, 102], ref["cell_source"][act])
    np.testing.assert_array_equal(got[act, 103], ref["cell_sink"][act])


@GRIDS
def test_advance_from_mf6_bitwise(tmp_path, removed):
    _, P, extras, _, _, _ = _read(str(tmp_path), removed)
    n = 600
    cell, lx, ly, lz = problems.release_in_block(P, n, (2, 7), (2, 5), (1, 3), seed=4)
    print("mf6 advance", n, "removed cells", len(removed), flush=True)
    E = _engine(P, extras)
    try:
        E.load_particles(cell, lx, ly, lz)
        st = O.new_state(cell, lx, ly, lz)
        OP = oracle_problem(P, extras)
        t = 0.0
        for k in range(1, 4):
            t1 = 0.5 * k
            E.advance(t1)
            assert O.advance(OP, st, t, t1) == 0
            t = t1
        g = E.particles()
    finally:
        E.close()
    for f in FIELDS:
        nan = f in ("lx", "ly", "lz", "t")
        assert np.array_equal(g[f], st[f], equal_nan=nan), (f, int((g[f] != st[f]).sum()))
    assert st["substeps"].sum() > n
    # the walk reached the statuses the boundary, sink and dry cells give
    assert len(set(st["status"].tolist())) >= 2


def test_driver_loads_each_time_step(tmp_path):
    """Two MODFLOW time steps with different flows: Driver with an
    mf6.FlowSeries gives the particles of the same set_flow,
    set_cell_extras and advance calls made by hand, and they differ from a
    run that keeps the first step's flow."""
    from rwpt.driver import Driver, schedule
    paths, P0, ibound = two_step_scenario(str(tmp_path))
    por, ret = np.full(P0.ncell, 0.25), np.ones(P0.ncell)
    fs = mf6.FlowSeries(*paths, steps=[(1, 1), (1, 2)], ibound=ibound, hdry=HDRY,
                        porosity=por, retardation=ret, default_iface=dict(IFACE),
                        aL=P0.aL, aT=P0.aT, dmeff=P0.dmeff, ts_params=(0.1, 0.0))
    ivs = schedule([1.0, 2.0], 2.0, time_points=[0.5, 1.5])
    assert [iv.ktime for iv in ivs] == [1, 1, 2, 2]
    rel = problems.release_in_block(P0, 400, (2, 7), (2, 5), (1, 3), seed=6)
    print("mf6 driver", len(ivs), flush=True)

    def run(mode):
        P1, x1 = fs(1)
        E = Engine(P1)
        try:
            E.set_cell_extras(**x1)
            E.load_particles(*rel)
            if mode == "driver":
                Driver(E, ivs, flow_for_step=fs).run()
            else:
                for iv in ivs:
                    if mode == "manual" and iv.ktime == 2 and iv.time == 1.5:
                        P2, x2 = fs(2)
                        E.set_flow(P2)
                        E.set_cell_extras(**x2)
                    E.advance(iv.time)
            return E.particles()
        finally:
            E.close()

    a, b, c = run("driver"), run("manual"), run("steady")
    for f in FIELDS:
        nan = f in ("lx", "ly", "lz", "t")
        assert np.array_equal(a[f], b[f], equal_nan=nan), f
    assert not np.array_equal(a["lx"], c["lx"], equal_nan=True)

