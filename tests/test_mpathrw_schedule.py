"""MODPATH-RW itself (MPathRW.F90 compiled from the reference's sources,
oracle/_ref/mpathrw) on a synthetic MODFLOW 6 run (tests/mpathrw_deck.py):
the particle arrays rwpt.driver.prepare_groups builds from the run's
particle groups equal what MPathRW reports for them in its endpoint file,
and the time points at which it writes timeseries records (every particle,
TimeseriesOutputOption 1) and gpkde reconstructions are those of
rwpt.driver.schedule, the dropped timeseries point included.  No device:
the engine's own files for the same run are compared byte for byte with
MPathRW's in tests/test_gpu_mpathrw.py."""
import numpy as np
import pytest

from oracle import refrun
from rwpt import driver

import mpathrw_deck as D

pytestmark = pytest.mark.skipif(not refrun.mpathrw_available(),
                                reason="oracle/_ref/mpathrw not built (needs the reference)")


def read_endpoints(path):
    lines = open(path).read().splitlines()
    body = lines[lines.index("END HEADER") + 1:]
    return np.array([[float(v) for v in ln.split()] for ln in body])


def read_records(path, header_end="END HEADER"):
    lines = open(path).read().splitlines()
    if header_end in lines:
        lines = lines[lines.index(header_end) + 1:]
    return [ln.split() for ln in lines if ln.strip()]


def groups_of(R):
    return [driver.ParticleGroup(G.name, G.release, G.cells, G.lx, G.ly, G.lz,
                                 offset=G.offset, ids=G.ids) for G in R.groups]


@pytest.fixture(scope="module", params=[(1, True), (1, False)],
                ids=["own_gpkde_points", "gpkde_follows_timeseries"])
def run(request, tmp_path_factory):
    ts_option, own = request.param
    R = D.default_run(timeseries_option=ts_option,
                      gpkde_points=(5.0, 10.0, 15.0) if own else None)
    tmp = str(tmp_path_factory.mktemp("mpathrw"))
    paths = D.write(R, tmp)
    refrun.mpathrw(paths["sim"])
    return R, paths


def test_particle_groups_as_mpathrw_prepares_them(run):
    R, paths = run
    g = driver.prepare_groups(groups_of(R))
    ep = read_endpoints(paths["out"]["mpend"])
    assert g["cell"].size == 40 + 24 + 3
    # the endpoint file lists the released particles (not status 0 or 8):
    # all but the three released after the stop time
    assert len(ep) == 64
    k = ep[:, 0].astype(int) - 1
    np.testing.assert_array_equal(k, np.arange(64))
    np.testing.assert_array_equal(ep[:, 1], g["group"][k])
    np.testing.assert_array_equal(ep[:, 2], g["id"][k])
    np.testing.assert_allclose(ep[:, 4], g["t"][k], rtol=1e-9, atol=0)
    np.testing.assert_array_equal(ep[:, 6], g["cell"][k])
    for c, f in ((8, "lx"), (9, "ly"), (10, "lz")):
        np.testing.assert_allclose(ep[:, c], g[f][k], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(ep[:, 15], g["face"][k])
    # later releases continue the IDs after the largest one read
    assert set(g["id"][40 + 12:40 + 24]) == set(range(113, 125))
    assert (g["face"] == 1).sum() == 2 and (g["face"] == 6).sum() == 22


def test_time_points_follow_the_schedule(run):
    R, paths = run
    own = R.gpkde_points is not None
    iv = driver.schedule(R.step_ends, R.stop_time, time_points=R.time_points,
                         gpkde_points=R.gpkde_points or (), gpkde_follows_timeseries=not own,
                         stopping_option=3, reconstruct=True)
    ts = read_records(paths["out"]["timeseries"])
    written = []
    for r in ts:
        key = (int(r[0]), float(r[2]))
        if not written or written[-1] != key:
            written.append(key)
    want = [(0, 0.0)] + [(i.time_point, i.time) for i in iv if i.time_point is not None]
    # every particle has a record at every point written (option 1)
    n = 40 + 24 + 3
    assert len(ts) == n * len(written)
    assert written == want[:len(written)] and len(written) >= 3
    if own:
        # the reconstruction at 5.0 ends the interval before the point at 6.0,
        # whose index is consumed: no record of point 2, nor of point 5 (17.5,
        # preceded by the reconstruction at 15.0)
        assert 2 not in [w[0] for w in written] and 5 not in [w[0] for w in written]
    kd = read_records(paths["out"]["gpkde.out"])
    blocks = []
    for r in kd:
        key = (int(r[0]), float(r[1]))
        if not blocks or blocks[-1] != key:
            blocks.append(key)
    want_kde = [(0, 0.0)] + [(i.gpkde_point, i.time) for i in iv if i.gpkde_point is not None]
    assert blocks == want_kde[:len(blocks)] and len(blocks) >= 2
