"""The unstructured-grid oracle against the reference itself.

oracle/ref/usg_ref.f90 runs the reference's own RectangularGridDisvMf6
(reading a MODFLOW 6 DISV binary grid file), ModpathCellData,
ParticleTrackingEngine neighbour fill, TrackCell and TrackSubCell,
compiled here from its Fortran sources.  The oracle's sub-cell tables
(connections, mass boundaries, velocities, sizes, offsets, time step,
corner discharge, porosity and dispersion) and its walks must equal the
reference's bit for bit, NaNs included.

Without the reference build, the oracle is still checked against the
structured oracle (itself pinned against the reference by
test_oracle_vs_reference.py) on a DISV grid with no refined cell, which is
a structured grid: both restatements must give the same walks.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import refrun
from rwpt import problems
from usg_scenarios import SCENARIOS, build, release

needs_ref = pytest.mark.skipif(not refrun.usg_available(),
                               reason="reference driver not built (oracle/ref)")


def _same(a, b):
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


@needs_ref
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_usg_tables_match_reference(name, tmp_path):
    Q, U = build(name)
    grb = str(tmp_path / "g.grb")
    Q.write_grb(grb)
    ref = refrun.usg_tables(grb, U)
    mine = O.usg_tables(U)
    assert len(mine) == sum(len(c) for c in ref)
    k = 0
    for n, subs in enumerate(ref):
        for ints, dbl in subs:
            o = mine[k]
            k += 1
            assert int(o[0]) == n + 1 and int(o[3]) == len(subs)
            got_i = np.r_[o[1:3], o[4:16], o[16:20]].astype(np.int32)
            assert np.array_equal(got_i, ints), (n + 1, got_i, ints)
            assert _same(o[20:], dbl).all(), (n + 1, ints[:2])


@needs_ref
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_usg_walk_matches_reference(name, tmp_path):
    Q, U = build(name)
    grb = str(tmp_path / "g.grb")
    Q.write_grb(grb)
    n = 96
    cell, lx, ly, lz = release(U, n)
    t_out = np.array([0.3, 0.7, 1.5])
    ref = refrun.usg_walk(grb, U, cell, lx, ly, lz, t_out)
    st = O.new_state(cell, lx, ly, lz)
    t0 = 0.0
    for k, t1 in enumerate(t_out):
        assert O.usg_advance(U, st, t0, t1) == 0
        t0 = t1
        for f in ("cell", "status", "lx", "ly", "lz", "t"):
            ok = _same(ref[k][f], st[f])
            assert ok.all(), (name, t1, f, np.nonzero(~ok)[0][:8])
    # the particles move (the stopping scenarios stop many at once)
    assert st["substeps"].sum() > 5 * n


def _structured_twin(Q, U):
    """The DIS problem of a DISV grid without refined cells."""
    nrow, ncol, nlay = Q.nrow, Q.ncol, Q.nlay
    ncpl = nrow * ncol
    frf = np.zeros(U.ncell)
    fff = np.zeros(U.ncell)
    flf = np.zeros(U.ncell)
    for c in range(U.ncell):
        for q in range(U.ia[c], U.ia[c + 1] - 1):
            m = U.ja[q] - 1
            if m == c + 1 and (c % ncol) + 1 < ncol:
                frf[c] = -U.flowja[q]
            elif m == c + ncol and (c % ncpl) // ncol + 1 < nrow:
                fff[c] = -U.flowja[q]
            elif m == c + ncpl:
                flf[c] = -U.flowja[q]
    return problems.Problem(
        ncol=ncol, nrow=nrow, nlay=nlay, delr=Q.del_x[:ncol].copy(),
        delc=Q.del_y[::ncol][:nrow].copy(), top=U.top, bot=U.bot, frf=frf, fff=fff,
        flf=flf, porosity=U.porosity, aL=U.aL, aT=U.aT, dmeff=U.dmeff,
        ts_params=U.ts_params, default_icbound=U.default_icbound)


@pytest.mark.parametrize("reflect", [0, 1])
def test_usg_oracle_equals_structured_oracle_without_refinement(reflect):
    Q, U = build("quad3")
    from usg_scenarios import GRIDS
    from usg_synth import Quadtree
    Q = Quadtree(**GRIDS["plain"])
    A = Q.cell_arrays()
    U = problems.UsgProblem(ncell=Q.ncell, nlay=Q.nlay, ia=Q.ia, ja=Q.ja,
                            flowja=Q.flows(seed=4), porosity=np.full(Q.ncell, 0.25),
                            aL=np.array([0.3]), aT=np.array([0.03]),
                            ts_params=(0.1, 0.0), default_icbound=reflect, **A)
    P = _structured_twin(Q, U)
    n = 200
    cell, lx, ly, lz = release(U, n, seed=9)
    a = O.new_state(cell, lx, ly, lz)
    b = O.new_state(cell, lx, ly, lz)
    t0 = 0.0
    for t1 in (0.4, 1.0, 2.5):
        O.usg_advance(U, a, t0, t1)
        O.advance(P, b, t0, t1)
        t0 = t1
        for f in ("cell", "status", "lx", "ly", "lz", "t", "substeps"):
            assert _same(a[f], b[f]).all(), (t1, f)
    assert (a["status"] == 2).any() or reflect
    assert a["substeps"].sum() > 20 * n
