"""The driver over the engine against MODPATH-RW itself.

MPathRW.F90, compiled from the reference's sources (oracle/_ref/mpathrw),
runs a synthetic MODFLOW 6 model (tests/mpathrw_deck.py: three time steps
in two stress periods with different flow rates, a convertible top layer,
a weak sink, three particle groups with offsets, a double release and
releases after the stop time, gpkde reconstructions at their own times or
at the timeseries points), on a DIS grid and on a DISV smoothed quadtree
(the unstructured engine, rwpt_create_usg).  rwpt.driver runs the same run on the device
from rwpt.mf6's reading of the same files: the timeseries and endpoint
files must be byte-identical to MPathRW's (the schedule with its dropped
timeseries points, the records at time 0, the release, stop and weak-sink
statuses, the sequence numbers and IDs of the groups, one interval after
the last particle stops), and every gpkde density row must match MPathRW's
(the same rows, bins and histogram; densities within the one-pass
reconstruction's summation-order tolerance, 1e-10 of the largest).  The
dispersion is zero, so the paths do not depend on either run's random
draws.
"""
import numpy as np
import pytest

from oracle import refrun
from rwpt import Engine, driver, engine, mf6

import mpathrw_deck as D

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refrun.mpathrw_available(),
                                 reason="oracle/_ref/mpathrw not built")]


def _engine_run(R, paths, tmp):
    f = paths["files"]
    g = mf6.read_grb(f["grb"])
    opts = dict(aL=np.array([0.0]), aT=np.array([0.0]), dmeff=np.array([0.0]),
                ts_params=(R.dt, 0.0), time_step_kind=4, advection_kind=2,
                stop_weak_sinks=1, default_icbound=0)
    fs = mf6.FlowSeries(f["grb"], f["hds"], f["bud"], steps=R.steps, ibound=None,
                        hdry=-1e30, porosity=R.porosity, default_iface={}, **opts)
    groups = [driver.ParticleGroup(G.name, G.release, G.cells, G.lx, G.ly, G.lz,
                                   offset=G.offset, ids=G.ids) for G in R.groups]
    pg = driver.prepare_groups(groups)
    own = R.gpkde_points is not None
    iv = driver.schedule(R.step_ends, R.stop_time, time_points=R.time_points,
                         gpkde_points=R.gpkde_points or (), gpkde_follows_timeseries=not own,
                         stopping_option=3, reconstruct=True)
    bin_size = (1.0, 1.0, 1.0)
    origin = (0.0, 0.0, float(np.min(g.botm)))
    nbins = (R.ncol, R.nrow, R.nlay)
    scale = 1.0 / np.prod(bin_size)
    kde = dict(origin=origin, bin_size=bin_size, nbins=nbins, mode=1, unit_volume=True,
               scaling_factor=scale, n_opt_loops=0, isotropic=False, rel_error=0.02)
    out = {k: str(tmp / f"engine.{k}") for k in ("timeseries", "mpend", "gpkde")}
    P1, x1 = fs(1)
    E = Engine(P1)
    try:
        if x1:
            E.set_cell_extras(**x1)
        E.load_particles(pg["cell"], pg["lx"], pg["ly"], pg["lz"], t=pg["t"],
                         status=np.zeros(pg["cell"].size, np.int32))
        engine.write_timeseries_header(out["timeseries"], 1, 0.0, (g.xorigin, g.yorigin),
                                       g.angrot)
        log = driver.Driver(
            E, iv, timeseries_path=out["timeseries"], sequence=pg["sequence"], ids=pg["id"],
            group=pg["group"], group_solute=pg["solute"], mass=pg["mass"], kde_options=kde,
            kde_path=out["gpkde"], flow_for_step=fs, timeseries_option=R.timeseries_option,
            kde_initial_options={"isotropic": False}, endpoint_path=out["mpend"],
            endpoint=dict(tracking_direction=1, reference_time=0.0,
                          origin=(g.xorigin, g.yorigin), rotation=g.angrot,
                          group_names=pg["names"], initial_face=pg["face"])).run()
        st = E.particles()["status"]
    finally:
        E.close()
    return out, log, st


def _rows(path):
    return [ln.split() for ln in open(path).read().splitlines() if ln.strip()]


@pytest.mark.parametrize("grid,ts_option,own",
                         [("dis", 0, True), ("dis", 1, True), ("dis", 0, False),
                          ("disv", 0, True), ("disv", 1, False)],
                         ids=["active_own_gpkde_points", "all_own_gpkde_points",
                              "active_gpkde_follows_timeseries", "disv_active_own_gpkde_points",
                              "disv_all_gpkde_follows_timeseries"])
def test_driver_files_equal_mpathrw(tmp_path, grid, ts_option, own):
    make = D.disv_run if grid == "disv" else D.default_run
    R = make(timeseries_option=ts_option, gpkde_points=(5.0, 10.0, 15.0) if own else None)
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    paths = D.write(R, str(ref_dir))
    print("mpathrw: run", grid, ts_option, own, flush=True)
    refrun.mpathrw(paths["sim"])
    print("mpathrw: engine", flush=True)
    out, log, st = _engine_run(R, paths, tmp_path)
    # the run reaches the statuses it is built for
    assert {0, 2, 3} <= set(st.tolist())
    ref_ts = open(paths["out"]["timeseries"], "rb").read()
    assert open(out["timeseries"], "rb").read() == ref_ts
    assert open(out["mpend"], "rb").read() == open(paths["out"]["mpend"], "rb").read()
    a, b = _rows(out["gpkde"]), _rows(paths["out"]["gpkde.out"])
    assert len(a) == len(b) and len(a) > 0
    ia = np.array([[int(v) for v in r[:1] + r[2:6]] for r in a])
    ib = np.array([[int(v) for v in r[:1] + r[2:6]] for r in b])
    np.testing.assert_array_equal(ia, ib)
    fa = np.array([[float(v) for v in (r[1],) + tuple(r[6:])] for r in a])
    fb = np.array([[float(v) for v in (r[1],) + tuple(r[6:])] for r in b])
    np.testing.assert_array_equal(fa[:, 0], fb[:, 0])           # times
    np.testing.assert_array_equal(fa[:, 2], fb[:, 2])           # histogram
    np.testing.assert_allclose(fa[:, 1], fb[:, 1], rtol=0, atol=1e-10 * fb[:, 1].max())
