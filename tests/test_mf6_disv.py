"""MODFLOW 6 DISV readers (rwpt.mf6 on a grid of rectangular cells)
against the reference's own readers.

A smoothed quadtree of three layers (tests/usg_synth.py) is written as a
DISV binary grid file, with head records of NCPL columns per layer and a
budget holding FLOW-JA-FACE, WEL and CHD lists (CHD with an IFACE
auxiliary variable), RCH with a default IFACE of 6, STO-SS storage and a
DATA-SPDIS record to skip.  rwpt.mf6 and the reference's
RectangularGridDisvMf6 (ReadGRB, CheckRectangular), HeadReader,
BudgetReader and FlowModelData%LoadTimeStep (oracle/_ref/mf6_ref with
"disv") read the same files: the cell geometry and FlowModelData's arrays
must agree bit for bit.  The engine's problem built from them
(mf6.usg_problem) is then checked through the oracle's unit records against
the reference's own assembly of the first sub-cell of every cell.
"""
import os

import numpy as np
import pytest

from rwpt import mf6

import mf6_synth as S
from usg_synth import Quadtree

HDRY = -1e30
IFACE = (("RCH", 6),)


def _refrun():
    from oracle import refrun
    return refrun


def scenario(tmp, seed=5):
    rng = np.random.default_rng(seed)
    Q = Quadtree(5, 6, {(0, 0), (1, 1), (2, 3), (3, 3), (4, 5)}, nlay=3)
    n, ncpl = Q.ncell, Q.ncpl
    icelltype = np.zeros(n, np.int32)
    icelltype[:ncpl] = 1
    grb, hds, cbc = (os.path.join(tmp, f) for f in ("m.disv.grb", "m.hds", "m.cbc"))
    Q.write_grb(grb, icelltype=icelltype)
    head = Q.botm + (Q.top - Q.botm) * rng.uniform(0.3, 1.4, n)
    dry = rng.choice(ncpl, 3, replace=False)
    head[dry] = HDRY
    S.write_heads(hds, Q.nlay, 1, ncpl, head)
    fja = Q.flows(seed=seed)
    wel = [ncpl + 3, 2 * ncpl + 7, ncpl + 3]
    rch = list(range(1, ncpl + 1, 4))
    chd = [k * ncpl + 1 for k in range(3)] + [k * ncpl + ncpl for k in range(3)]
    recs = [
        ("list", "WEL", "WEL-1", wel, [-5.0, -2.5, 3.0], (), None),
        ("list", "RCH", "RCH-1", rch, list(rng.uniform(0.01, 0.2, len(rch))), (), None),
        ("list", "CHD", "CHD-1", chd, list(rng.uniform(-3, 3, len(chd))), ("IFACE",),
         [[1.0 if m % 2 else 2.0] for m in range(len(chd))]),
        ("array", "STO-SS", np.where(rng.uniform(size=n) < 0.2, rng.normal(size=n), 0.0)),
        ("list", "DATA-SPDIS", "MODEL", list(range(1, n + 1)), [0.0] * n,
         ("QX", "QY", "QZ"), rng.normal(size=(n, 3)).tolist()),
    ]
    S.write_budget(cbc, Q.nlay, 1, ncpl, fja, recs)
    return Q, (grb, hds, cbc)


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    R = _refrun()
    if not R.mf6_available():
        pytest.skip("oracle/_ref/mf6_ref is not built (needs the reference sources)")
    tmp = str(tmp_path_factory.mktemp("disv"))
    Q, paths = scenario(tmp)
    n = Q.ncell
    ibound = np.ones(n, np.int32)
    por = np.full(n, 0.3)
    ret = np.ones(n)
    ref = R.mf6(*paths, ibound, por, ret, hdry=HDRY, default_iface=IFACE, disv=True)
    g = mf6.read_grb(paths[0])
    ts = mf6.load_time_step(g, mf6.read_heads(paths[1]), mf6.read_budget(paths[2]), 1, 1,
                            ibound, HDRY, dict(IFACE))
    return dict(Q=Q, paths=paths, ref=ref, grb=g, ts=ts, ibound=ibound, por=por)


def test_disv_grid_file(case):
    g, ref, Q = case["grb"], case["ref"], case["Q"]
    assert isinstance(g, mf6.GrbDisv)
    assert (g.ncells, g.nlay, g.ncpl) == (Q.ncell, Q.nlay, Q.ncpl)
    for k in ("cell_x", "cell_y", "del_x", "del_y"):
        np.testing.assert_array_equal(getattr(g, k), ref[k], err_msg=k)
    np.testing.assert_array_equal(g.top, ref["top"])
    np.testing.assert_array_equal(g.botm, ref["bottom"])
    # the vertices give back the generator's cells
    A = Q.cell_arrays()
    np.testing.assert_array_equal(g.del_x, A["del_x"])
    np.testing.assert_array_equal(g.cell_y, A["cell_y"])
    np.testing.assert_array_equal(g.layer, A["layer"])


def test_disv_time_step_arrays_bitwise(case):
    ts, ref = case["ts"], case["ref"]
    np.testing.assert_array_equal(ts.ibound_ts, ref["ibound_ts"])
    for k in ("heads", "flows_ja", "boundary_flows", "source_flows", "sink_flows",
              "storage_flows"):
        np.testing.assert_array_equal(getattr(ts, k), ref[k], err_msg=k)
    assert ts.steady_state == ref["steady_state"]
    np.testing.assert_array_equal(ts.saturated_top, ref["saturated_top"])
    assert (ts.saturated_top == ts.heads).sum() >= 3
    assert ((ts.saturated_top == case["grb"].botm) & (ts.ibound_ts == 0)).sum() >= 3
    assert (ts.ibound_ts == 0).sum() >= 3
    assert np.count_nonzero(ts.boundary_flows) >= 5


def test_disv_problem_assembles_like_the_reference(case):
    """The oracle's unit records from mf6.usg_problem against the reference's
    SetMassTransportDataUnstructured and FillMassSubCellDataBuffer of
    sub-cell (1, 1) of every cell: face pore velocities, sizes and the
    cell's source and sink flows."""
    from oracle import oracle as O
    U = mf6.usg_problem(case["grb"], case["ts"], case["por"])
    T = O.usg_tables(U)
    first = np.concatenate([[True], T[1:, 0] != T[:-1, 0]])
    T1 = T[first]
    ref = case["ref"]
    assert len(T1) == U.ncell
    np.testing.assert_array_equal(T1[:, 20:26], ref["velocity"])
    np.testing.assert_array_equal(T1[:, 26:29], ref["size"])
    np.testing.assert_array_equal(T1[:, 36], ref["cell_source"])
    np.testing.assert_array_equal(T1[:, 37], ref["cell_sink"])
    assert (T[:, 3] == 4).any()
