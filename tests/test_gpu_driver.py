"""The driver over the engine: advancing through the schedule writes the
timeseries records of the active particles at every timeseries point (and
none at the points the schedule drops) and one block of density rows per
reconstruction, identical to the same calls made by hand."""
import numpy as np
import pytest

from rwpt import Engine, engine, problems
from rwpt.driver import Driver, schedule

pytestmark = pytest.mark.gpu


def test_driver_outputs(tmp_path):
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    rel = problems.release_in_block(P, 2000, (2, 6), (6, 15), (1, 6), seed=3)
    iv = schedule([1e30], 6.0, time_points=[1.0, 2.0, 4.0], gpkde_points=[1.5, 2.0, 5.0],
                  gpkde_follows_timeseries=False, reconstruct=True)
    kde = {"origin": (0.0, 0.0, float(P.bot.min())), "bin_size": (1.0, 1.0, 1.0),
           "nbins": (P.ncol, P.nrow, P.nlay), "mode": 1}
    ts, kd = tmp_path / "ts.out", tmp_path / "kde.out"
    ts2, kd2 = tmp_path / "ts2.out", tmp_path / "kde2.out"
    print("driver: intervals", [(i.time, i.time_point, i.gpkde_point) for i in iv], flush=True)
    E = Engine(P)
    try:
        E.load_particles(*rel)
        engine.write_timeseries_header(str(ts), 1, 0.0, (0.0, 0.0), 0.0)
        log = Driver(E, iv, timeseries_path=str(ts), kde_options=kde,
                     kde_path=str(kd), initial_point=False).run()
    finally:
        E.close()
    assert [x[0].time for x in log] == [1.0, 1.5, 2.0, 5.0, 6.0]
    # the same calls by hand
    seq = np.arange(1, 2001, dtype=np.int32)
    E = Engine(P)
    try:
        E.load_particles(*rel)
        engine.write_timeseries_header(str(ts2), 1, 0.0, (0.0, 0.0), 0.0)
        first = True
        for t, tp, gp in [(1.0, 1, None), (1.5, None, 1), (2.0, None, 2), (5.0, None, 3),
                          (6.0, None, None)]:
            E.advance(t)
            if tp is not None:
                E.write_timeseries(str(ts2), tp, 1, t, seq, seq, np.ones(2000, np.int32))
            if gp is not None:
                d, h = E.kde(mass=np.ones(2000), **kde)
                engine.kde_write(str(kd2), kde["origin"], kde["bin_size"], kde["nbins"], d, h,
                                 time_id=gp, time=t, append=not first)
                first = False
    finally:
        E.close()
    assert ts.read_bytes() == ts2.read_bytes()
    assert kd.read_bytes() == kd2.read_bytes()
    lines = ts.read_text().splitlines()
    assert lines[2] == "END HEADER" and len(lines) > 3
    assert {ln.split()[0] for ln in lines[3:]} == {"1"}   # only point 1 is written
