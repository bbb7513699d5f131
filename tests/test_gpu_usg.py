"""GPU parity on unstructured grids of rectangular cells (MODFLOW 6 DISV
smoothed quadtrees): the HIP unit records and walk (rwpt_create_usg,
rwpt_set_flow_usg, rwpt_advance through the C ABI) against the oracle's
unstructured restatement, which test_oracle_usg_vs_reference.py pins bit
for bit against the reference compiled from its sources.

EULERIAN runs agree bit for bit (cell, status, local coordinates, time,
sub-steps and rng state after every interval, NaNs included); the
EXPONENTIAL scenario calls exp, whose last bit differs between the device
and host math libraries, and is compared with a tolerance.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rwpt import Engine, problems
from usg_scenarios import GRIDS, SCENARIOS, build, release
from usg_synth import Quadtree

pytestmark = pytest.mark.gpu

FIELDS = ("cell", "status", "lx", "ly", "lz", "t", "substeps", "rng")
T_OUT = (0.3, 0.7, 1.5)


def _same(a, b):
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_usg_tables_bitwise(name):
    Q, U = build(name)
    E = Engine(U)
    try:
        assert E.unit_count() == len(O.usg_tables(U))
        got = E.usg_tables()
    finally:
        E.close()
    want = O.usg_tables(U)
    ok = _same(got, want)
    bad = np.nonzero(~ok.all(1))[0]
    assert ok.all(), (name, bad[:5], np.nonzero(~ok[bad[0]])[0] if bad.size else None)


def _run_both(U, n, t_out=T_OUT, seed=3, tuning=None):
    cell, lx, ly, lz = release(U, n, seed=seed)
    E = Engine(U)
    try:
        if tuning:
            E.set_tuning(**tuning)
        E.load_particles(cell, lx, ly, lz)
        st = O.new_state(cell, lx, ly, lz)
        t0 = 0.0
        out = []
        for t1 in t_out:
            act, ns = E.advance(t1)
            before = st["substeps"].sum()
            assert O.usg_advance(U, st, t0, t1) == 0
            t0 = t1
            g = E.particles()
            out.append((g, {f: st[f].copy() for f in FIELDS}, act, ns,
                         st["substeps"].sum() - before))
    finally:
        E.close()
    return out


@pytest.mark.parametrize("name", sorted(n for n in SCENARIOS if "exponential" not in n))
def test_usg_advance_bitwise(name):
    Q, U = build(name)
    for k, (g, st, act, ns, ns_orc) in enumerate(_run_both(U, 400)):
        assert ns == ns_orc and act == int((st["status"] == 1).sum()), (name, k)
        for f in FIELDS:
            ok = _same(g[f], st[f])
            assert ok.all(), (name, T_OUT[k], f, np.nonzero(~ok)[0][:8])
    # the scenarios stop particles in more than one way
    assert len(np.unique(st["status"])) >= 2


def test_usg_advance_exponential_close():
    Q, U = build("quad3_exponential")
    g, st, _, _, _ = _run_both(U, 400)[-1]
    same = (g["cell"] == st["cell"]) & (g["status"] == st["status"])
    assert same.mean() >= 0.99
    for f in ("lx", "ly", "lz", "t"):
        a, b = g[f][same], st[f][same]
        fin = np.isfinite(a) & np.isfinite(b)
        assert np.allclose(a[fin], b[fin], rtol=1e-10, atol=1e-10), f


@pytest.mark.parametrize("tuning", [dict(budget=1, block=64, sort=False),
                                    dict(budget=7, block=128, sort=True),
                                    dict(budget=1000, block=256, sort=True)])
def test_usg_launch_shape_does_not_change_results(tuning):
    Q, U = build("checker")
    ref = _run_both(U, 300)
    got = _run_both(U, 300, tuning=tuning)
    for (a, *_), (b, *_) in zip(ref, got):
        for f in FIELDS:
            assert _same(a[f], b[f]).all(), f


def test_usg_larger_grid_bitwise():
    """A 24 x 30 base grid with a refined block and a refined ring, three
    layers, 20k particles, four intervals."""
    refine = {(i, j) for i in range(8, 16) for j in range(10, 20)}
    refine |= {(i, j) for i in range(2, 22) for j in (3, 26)}
    Q = Quadtree(24, 30, refine, nlay=3)
    A = Q.cell_arrays()
    U = problems.UsgProblem(ncell=Q.ncell, nlay=Q.nlay, ia=Q.ia, ja=Q.ja,
                            flowja=Q.flows(seed=21), porosity=np.full(Q.ncell, 0.3),
                            aL=np.array([0.5]), aT=np.array([0.05]), ts_params=(0.1, 0.0),
                            **A)
    out = _run_both(U, 20000, t_out=(1.0, 2.0, 4.0, 8.0), seed=5)
    for k, (g, st, act, ns, ns_orc) in enumerate(out):
        assert ns == ns_orc
        for f in FIELDS:
            ok = _same(g[f], st[f])
            assert ok.all(), (k, f, np.nonzero(~ok)[0][:8])
    assert out[-1][3] > 20000 * 20


def test_usg_global_xyz_and_histogram():
    Q, U = build("block")
    # a saturated top below the cell top in the upper layer
    U.saturated_top = U.top.copy()
    U.saturated_top[:Q.ncpl] = U.bot[:Q.ncpl] + 0.6 * (U.top - U.bot)[:Q.ncpl]
    cell, lx, ly, lz = release(U, 2000, seed=8)
    E = Engine(U)
    try:
        E.load_particles(cell, lx, ly, lz)
        E.advance(0.5)
        g = E.particles()
        gx, gy, gz = E.global_xyz()
        origin, binsz, nb = (0.0, 0.0, 0.0), (0.5, 0.5, 1.0), (12, 12, 2)
        dens, hist = E.kde(origin, binsz, nb, mode=0)
    finally:
        E.close()
    ox, oy, oz = O.usg_global_xyz(U, g["cell"], g["lx"], g["ly"], g["lz"])
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy) and np.array_equal(gz, oz)
    act = g["status"] == 1
    want = O.histogram(ox[act], oy[act], oz[act], origin, binsz, nb)
    assert np.array_equal(hist, want)


def test_usg_plain_grid_equals_structured_engine(tmp_path):
    """A DISV grid without refined cells is a structured grid: the
    unstructured walk must give the structured walk's results, and the
    timeseries and observation records written from both engines (global
    coordinates with a saturated top, water volumes) the same bytes."""
    from test_oracle_usg_vs_reference import _structured_twin
    Q = Quadtree(**GRIDS["plain"])
    A = Q.cell_arrays()
    U = problems.UsgProblem(ncell=Q.ncell, nlay=Q.nlay, ia=Q.ia, ja=Q.ja,
                            flowja=Q.flows(seed=4), porosity=np.full(Q.ncell, 0.25),
                            aL=np.array([0.3]), aT=np.array([0.03]), ts_params=(0.1, 0.0),
                            **A)
    P = _structured_twin(Q, U)
    sat = U.top.copy()
    sat[:Q.ncpl] -= 0.3 * (U.top - U.bot)[:Q.ncpl]
    U.saturated_top = P.saturated_top = sat
    cell, lx, ly, lz = release(U, 1000, seed=9)
    n = cell.size
    ids, group, solute = np.arange(1, n + 1, dtype=np.int32), np.ones(n, np.int32), \
        np.ones(1, np.int32)
    obs_cells = np.arange(1, Q.ncell + 1, dtype=np.int32)
    res, files = [], []
    for k, prob in enumerate((U, P)):
        E = Engine(prob)
        try:
            E.load_particles(cell, lx, ly, lz)
            for t1 in (0.4, 1.0, 2.5):
                E.advance(t1)
            res.append(E.particles())
            ts, ob = tmp_path / f"ts{k}", tmp_path / f"obs{k}"
            E.write_timeseries(str(ts), 3, 1, 2.5, ids, ids, group)
            E.write_observation(str(ob), 1, obs_cells, 3, 1, ids, group, solute,
                                fmt=0, append=False)
            files.append((ts.read_bytes(), ob.read_bytes()))
        finally:
            E.close()
    for f in FIELDS:
        assert _same(res[0][f], res[1][f]).all(), f
    assert files[0][0] == files[1][0] and len(files[0][0]) > 0
    assert files[0][1] == files[1][1] and len(files[0][1]) > 0


def test_usg_from_mf6_disv_files(tmp_path):
    """A DISV model read from its binary grid, head and budget files
    (rwpt.mf6, pinned against the reference's readers by test_mf6_disv.py):
    the engine's walk over mf6.usg_problem equals the oracle's."""
    from rwpt import mf6
    from test_mf6_disv import HDRY, IFACE, scenario
    Q, paths = scenario(str(tmp_path))
    g = mf6.read_grb(paths[0])
    ts = mf6.load_time_step(g, mf6.read_heads(paths[1]), mf6.read_budget(paths[2]), 1, 1,
                            np.ones(g.ncells, np.int32), HDRY, dict(IFACE))
    U = mf6.usg_problem(g, ts, 0.3, aL=np.array([0.3]), aT=np.array([0.03]),
                        ts_params=(0.1, 0.0))
    out = _run_both(U, 600, t_out=(0.5, 1.0, 2.0), seed=4)
    for k, (gp, st, act, ns, ns_orc) in enumerate(out):
        assert ns == ns_orc
        for f in FIELDS:
            ok = _same(gp[f], st[f])
            assert ok.all(), (k, f, np.nonzero(~ok)[0][:8])
    # dry cells (IBoundTS 0) stop the particles released in them
    assert (out[-1][1]["status"] == 7).any()


def test_usg_autotune_keeps_results():
    """rwpt_autotune on an unstructured engine: its trial advances restore the
    particles (the unstructured visit state included), so the advances after
    tuning equal those of an untuned engine bit for bit."""
    Q, U = build("checker")
    cell, lx, ly, lz = release(U, 600, seed=3)
    res = []
    for tune in (False, True):
        E = Engine(U)
        try:
            E.load_particles(cell, lx, ly, lz)
            E.advance(0.3)
            if tune:
                best = E.autotune(0.7)
                assert best["block"] in (64, 128, 256)
            E.advance(0.7)
            E.advance(1.5)
            res.append(E.particles())
        finally:
            E.close()
    for f in FIELDS:
        assert _same(res[0][f], res[1][f]).all(), f
