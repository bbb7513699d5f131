"""MODFLOW 6 DIS readers (rwpt.mf6) against the reference's own readers.

No MODFLOW 6 output ships with the reference, so each case writes a small
DIS simulation's binary grid, head and budget files (tests/mf6_synth.py)
and reads them twice: with rwpt.mf6 and with the reference's
RectangularGridDisMf6, HeadReader, BudgetReader and
FlowModelData%LoadTimeStep compiled from its sources (oracle/_ref/mf6_ref).
FlowModelData's arrays must agree bit for bit.  The cells the engine builds
from the reader's face flows are then compared with the reference's
MODFLOW 6 assembly (SetMassTransportDataUnstructured and
FillMassSubCellDataBuffer): face pore velocities, sizes and the cell's
source and sink flows, through the C oracle's cell records here and the
engine's own on the GPU (tests/test_gpu_mf6.py).

The budget holds FLOW-JA-FACE, WEL and RIV lists (source/sink by sign),
RCH with a default IFACE of 6 (top face), CHD and RIV with an IFACE
auxiliary variable (boundary faces, and 0 = source/sink), STO-SS storage
and a DATA-SPDIS record that must be skipped.  Layer 1 is convertible
(ICELLTYPE 1) with some dry cells (head = HDRY) and a head below a bottom.
Each test runs on that grid and on the same grid with six cells removed
(IDOMAIN 0: empty rows in IA/JA, no flow across their faces, head
HNOFLOW), whose IBOUND then comes from IDOMAIN as in ModpathBasicData.
"""
import os

import numpy as np
import pytest

from rwpt import mf6, problems

import mf6_synth as S

HERE = os.path.dirname(os.path.abspath(__file__))


def _refrun():
    from oracle import refrun
    return refrun


HDRY = -1e30
IFACE = (("RCH", 6), ("EVT", 6))


# cells removed with IDOMAIN 0 (k, i, j, 1-based) in the "idomain" case: a
# block through two layers, a column end and a single cell, none of them
# holding a boundary package
REMOVED = [(2, 3, 2), (2, 3, 3), (3, 3, 2), (3, 3, 3), (1, 5, 8), (3, 1, 5)]


def scenario(tmp, seed=3, nlay=3, nrow=6, ncol=8, removed=()):
    """Writes the three files; returns (paths, problem, ibound).  Cells in
    `removed` get IDOMAIN 0: no connections in the grid file, head HNOFLOW
    (1e30), no flow across their faces (the problem's face flows there are
    set to 0 to match) and IBOUND 0."""
    rng = np.random.default_rng(seed)
    P = problems.heterogeneous_flow(ncol, nrow, nlay, corr_len=2.0, seed=seed)
    n, ncpl = P.ncell, nrow * ncol
    idomain = np.ones(n, np.int32)
    for k, i, j in removed:
        c = int(P.cell_number(k, i, j)) - 1
        idomain[c] = 0
        # faces of c: right/left (frf), front/back (fff), lower/upper (flf)
        P.frf[c] = 0.0
        if c % ncol > 0:
            P.frf[c - 1] = 0.0
        P.fff[c] = 0.0
        if (c % ncpl) // ncol > 0:
            P.fff[c - ncol] = 0.0
        P.flf[c] = 0.0
        if c >= ncpl:
            P.flf[c - ncpl] = 0.0
    icelltype = np.zeros(n, np.int32)
    icelltype[:ncpl] = 1
    head = P.bot + (P.top - P.bot) * rng.uniform(0.3, 1.4, n)
    dry = rng.choice(ncpl, 4, replace=False)
    head[dry] = HDRY
    head[dry[0] + 1 if dry[0] + 1 not in dry else dry[0]] = P.bot[0] - 1.0
    head[idomain == 0] = 1e30
    grb, hds, cbc = (os.path.join(tmp, f) for f in ("m.dis.grb", "m.hds", "m.cbc"))
    ia, ja = S.write_grb(grb, nlay, nrow, ncol, P.delr, P.delc, P.top[:ncpl], P.bot,
                         icelltype, idomain)
    S.write_heads(hds, nlay, nrow, ncol, head)
    fja = S.flow_ja(ia, ja, P.frf, P.fff, P.flf, nrow, ncol)
    cells = lambda k, i, j: int(P.cell_number(k, i, j))
    wel = [cells(2, 3, 4), cells(3, 5, 6), cells(2, 3, 4)]
    rch = [cells(1, i, j) for i in (1, 2, 6) for j in (1, 5, 8)]
    chd = [cells(k, i, 1) for k in (1, 2, 3) for i in (2, 4)]
    riv = [cells(1, 3, 3), cells(1, 4, 7), cells(2, 1, 8)]
    recs = [
        ("list", "WEL", "WEL-1", wel, [-5.0, -2.5, 3.0], (), None),
        ("list", "RCH", "RCH-1", rch, list(rng.uniform(0.01, 0.2, len(rch))), (), None),
        ("list", "CHD", "CHD-1", chd, list(rng.uniform(-3, 3, len(chd))), ("IFACE",),
         [[1.0 if m % 3 else 0.0] for m in range(len(chd))]),
        ("list", "RIV", "RIV-1", riv, [-1.5, 0.75, -0.25], ("IFACE", "COND"),
         [[6.0, 10.0], [0.0, 3.0], [2.0, 1.0]]),
        ("array", "STO-SS", np.where((rng.uniform(size=n) < 0.2) & (idomain == 1),
                                     rng.normal(size=n), 0.0)),
        ("list", "DATA-SPDIS", "MODEL", list(range(1, n + 1)), [0.0] * n,
         ("QX", "QY", "QZ"), rng.normal(size=(n, 3)).tolist()),
    ]
    S.write_budget(cbc, nlay, nrow, ncol, fja, recs)
    if removed:
        # MODPATH's IBOUND of a MODFLOW 6 grid comes from IDOMAIN
        ibound = mf6.ibound_from_idomain(mf6.read_grb(grb))
    else:
        ibound = np.ones(n, np.int32)
        ibound[cells(3, 6, 8) - 1] = 0
    return (grb, hds, cbc), P, ibound


@pytest.fixture(scope="module", params=["full", "idomain"])
def case(request, tmp_path_factory):
    R = _refrun()
    if not R.mf6_available():
        pytest.skip("oracle/_ref/mf6_ref is not built (needs the reference sources)")
    tmp = str(tmp_path_factory.mktemp("mf6"))
    removed = REMOVED if request.param == "idomain" else ()
    paths, P, ibound = scenario(tmp, removed=removed)
    por = np.full(P.ncell, 0.25)
    por[::5] = 0.3
    ret = np.ones(P.ncell)
    ret[::7] = 1.5
    ref = R.mf6(*paths, ibound, por, ret, hdry=HDRY, default_iface=IFACE)
    g = mf6.read_grb(paths[0])
    ts = mf6.load_time_step(g, mf6.read_heads(paths[1]), mf6.read_budget(paths[2]), 1, 1,
                            ibound, HDRY, dict(IFACE))
    return dict(paths=paths, P=P, ibound=ibound, por=por, ret=ret, ref=ref, grb=g, ts=ts,
                removed=removed)


def test_grid_file(case):
    g, ref, P = case["grb"], case["ref"], case["P"]
    assert (g.nlay, g.nrow, g.ncol) == (P.nlay, P.nrow, P.ncol)
    np.testing.assert_array_equal(g.top, ref["top"])
    np.testing.assert_array_equal(g.botm, ref["bottom"])
    np.testing.assert_array_equal(g.top, P.top)
    assert g.nja == len(ref["flows_ja"])
    assert (g.idomain == 0).sum() == len(case["removed"])
    np.testing.assert_array_equal(mf6.ibound_from_idomain(g) == 0, g.idomain < 1)


def test_time_step_arrays_bitwise(case):
    ts, ref = case["ts"], case["ref"]
    np.testing.assert_array_equal(ts.ibound_ts, ref["ibound_ts"])
    for k in ("heads", "flows_ja", "boundary_flows", "source_flows", "sink_flows",
              "storage_flows"):
        np.testing.assert_array_equal(getattr(ts, k), ref[k], err_msg=k)
    assert ts.steady_state == ref["steady_state"]
    # SaturatedTop: heads of wet convertible cells, bottoms of dry ones
    np.testing.assert_array_equal(ts.saturated_top, ref["saturated_top"])
    g = case["grb"]
    assert (ts.saturated_top == ts.heads).sum() >= 3
    assert ((ts.saturated_top == g.botm) & (ts.ibound_ts == 0)).sum() >= 2
    # the scenario exercises what it claims
    assert (ts.ibound_ts == 0).sum() >= 4
    assert np.count_nonzero(ts.boundary_flows) >= 5
    assert np.count_nonzero(ts.source_flows) >= 1 and np.count_nonzero(ts.sink_flows) >= 2
    assert not ts.steady_state


def test_problem_carries_the_steps_steady_state(case):
    """FlowModelData%SteadyState of the read step goes to the engine's
    Problem (the Driver hands it to rwpt_set_steady_state)."""
    P, _ = mf6.problem(case["grb"], case["ts"], case["por"], case["ret"])
    assert P.steady_state == int(case["ref"]["steady_state"]) == 0


def test_face_flows_invert_flow_ja(case):
    """The MODFLOW-2005 face flows of the reader are the ones the files were
    written from."""
    a, P = mf6.engine_arrays(case["grb"], case["ts"]), case["P"]
    for k in ("frf", "fff", "flf"):
        np.testing.assert_array_equal(a[k], getattr(P, k), err_msg=k)


def engine_problem(P0, grb, ts, por, ret):
    """The engine's Problem and cell extras from one read time step."""
    import dataclasses
    a = mf6.engine_arrays(grb, ts)
    P = dataclasses.replace(P0, frf=a["frf"], fff=a["fff"], flf=a["flf"], bflow=a["bflow"],
                            porosity=por, retardation=ret, ibound=a["ibound"])
    extras = dict(head=a["head"], layer_type=a["layer_type"], source_flow=a["source_flow"],
                  sink_flow=a["sink_flow"])
    return P, extras


def oracle_problem(P, extras):
    """The C oracle's problem with the cell extras set."""
    from oracle import oracle as O
    OP = O.OracleProblem(P)
    keep = {k: np.ascontiguousarray(extras[k], np.float64) for k in
            ("head", "source_flow", "sink_flow")}
    keep["layer_type"] = np.ascontiguousarray(extras["layer_type"], np.int32)
    OP.s.head = O._ptr(keep["head"], O._dp)
    OP.s.source_flow = O._ptr(keep["source_flow"], O._dp)
    OP.s.sink_flow = O._ptr(keep["sink_flow"], O._dp)
    OP.s.layer_type = O._ptr(keep["layer_type"], O._ip)
    OP._keep.update(keep)
    return OP


def _oracle_tables(case):
    from oracle import oracle as O
    P, extras = engine_problem(case["P"], case["grb"], case["ts"], case["por"], case["ret"])
    return O.cell_tables(oracle_problem(P, extras))


def test_cells_match_reference_mf6_assembly(case):
    """The cells built from the reader's arrays by the structured assembly
    (the C oracle's, which the engine's matches bit for bit) equal the
    reference's MODFLOW 6 assembly: face velocities, DX DY DZ and the
    source and sink flows after boundary flows on faces without an active
    neighbour were folded in."""
    tab, ref = _oracle_tables(case), case["ref"]
    act = case["ts"].ibound_ts != 0
    np.testing.assert_array_equal(tab[act, 0:6], ref["velocity"][act])
    np.testing.assert_array_equal(tab[act, 6:9], ref["size"][act])
    np.testing.assert_array_equal(tab[act, 102], ref["cell_source"][act])
    np.testing.assert_array_equal(tab[act, 103], ref["cell_sink"][act])
    assert np.count_nonzero(ref["cell_source"][act]) >= 1


def test_rejects_other_files(tmp_path):
    p = tmp_path / "x.grb"
    p.write_bytes(b"GRID DISV".ljust(50) + b"\0" * 200)
    with pytest.raises(mf6.Mf6Error):
        mf6.read_grb(str(p))
    paths, P, ibound = scenario(str(tmp_path))
    g = mf6.read_grb(paths[0])
    g.idomain[0] = 0
    ts = mf6.load_time_step(g, mf6.read_heads(paths[1]), mf6.read_budget(paths[2]), 1, 1,
                            ibound, HDRY)
    with pytest.raises(mf6.Mf6Error):
        # a removed cell whose row in JA is not empty
        mf6.engine_arrays(g, ts)
    with pytest.raises(mf6.Mf6Error):
        mf6.load_time_step(mf6.read_grb(paths[0]), mf6.read_heads(paths[1]),
                           mf6.read_budget(paths[2]), 2, 1, ibound, HDRY)


def two_step_scenario(tmp):
    """scenario() with a second time step (kper 1, kstp 2): heads raised by
    0.125 outside the dry cells, face flows doubled, other wells and
    recharge, no storage."""
    paths, P, ibound = scenario(tmp)
    h1 = np.concatenate([r.data for r in mf6.read_heads(paths[1])])
    h2 = np.where(h1 == HDRY, HDRY, h1 + 0.125)
    S.write_heads(paths[1], P.nlay, P.nrow, P.ncol, [h1, h2], steps=((1, 1, 1.0), (1, 2, 2.0)))
    g = mf6.read_grb(paths[0])
    fja2 = 2.0 * S.flow_ja(g.ia + 1, g.ja, P.frf, P.fff, P.flf, P.nrow, P.ncol)
    wel = [int(P.cell_number(2, 3, 4)), int(P.cell_number(1, 2, 2))]
    S.write_budget(paths[2], P.nlay, P.nrow, P.ncol, fja2,
                   [("list", "WEL", "WEL-1", wel, [-4.0, 1.0], (), None),
                    ("list", "RCH", "RCH-1", [1, 2, 3], [0.5, 0.25, 0.125], (), None)],
                   kper=1, kstp=2, totim=2.0, append=True)
    return paths, P, ibound


def test_second_time_step(tmp_path):
    """Files with two time steps: the reader keeps only the requested one
    (the other is skipped unread) and equals the reference's
    LoadTimeStep(1, 2)."""
    R = _refrun()
    if not R.mf6_available():
        pytest.skip("oracle/_ref/mf6_ref is not built (needs the reference sources)")
    paths, P, ibound = two_step_scenario(str(tmp_path))
    g = mf6.read_grb(paths[0])
    por, ret = np.full(P.ncell, 0.25), np.ones(P.ncell)
    ref = R.mf6(*paths, ibound, por, ret, kper=1, kstp=2, hdry=HDRY, default_iface=IFACE)
    heads = mf6.read_heads(paths[1], kper=1, kstp=2)
    budget = mf6.read_budget(paths[2], kper=1, kstp=2)
    assert {(r.kper, r.kstp) for r in heads + budget} == {(1, 2)}
    ts = mf6.load_time_step(g, heads, budget, 1, 2, ibound, HDRY, dict(IFACE))
    np.testing.assert_array_equal(ts.ibound_ts, ref["ibound_ts"])
    for k in ("heads", "flows_ja", "boundary_flows", "source_flows", "sink_flows",
              "storage_flows"):
        np.testing.assert_array_equal(getattr(ts, k), ref[k], err_msg=k)
    assert ts.steady_state == ref["steady_state"] and ts.steady_state
    np.testing.assert_array_equal(mf6.engine_arrays(g, ts)["frf"], 2.0 * P.frf)

