"""The density-row writer (rwpt_kde_write) against the file the reference's
GridProjectedKDE writes itself (ExportDensityUnit / ExportDensityUnitBinary,
GridProjectedKDE.F90:6504-7021, with the time index, time and particle
group MPathRW.F90:1626-1642 passes).  The writer is fed the reference's own
density and histogram, so the files must be byte-identical.  Host code
only: runs without a GPU."""
import numpy as np
import pytest

from oracle import refrun
from rwpt import engine

pytestmark = pytest.mark.skipif(not refrun.gpkde_available(),
                                reason="compiled reference gpkde not built")


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("col", [0, 1, 2])
@pytest.mark.parametrize("case", ["2d", "3d_w"])
def test_rows_match_reference_file(tmp_path, fmt, col, case):
    g = np.random.default_rng(3)
    n = 3000
    if case == "2d":
        nb, bs, origin = (40, 30, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
        pts = (g.normal(20, 4, n), g.normal(15, 3, n), g.uniform(0, 1, n))
        w = None
    else:
        nb, bs, origin = (20, 16, 12), (0.5, 2.0, 1.5), (-3.0, 10.0, 2.5)
        pts = (g.normal(2, 2, n), g.normal(26, 6, n), g.normal(11.5, 3, n))
        w = g.uniform(0.2, 3.0, n)
    sf = 0.37
    tid, tval, gid = 7, 123.456789, 3
    d_ref, c_ref, _, blob = refrun.gpkde(*pts, origin, bs, nb, weights=w, unit_volume=True,
                                         scaling_factor=sf,
                                         export=(fmt, col, tid, tval, gid))
    out = tmp_path / "rows.out"
    engine.kde_write(str(out), origin, bs, nb, d_ref, c_ref, time_id=tid, time=tval,
                     group_id=gid, fmt=fmt, column_format=col)
    mine = out.read_bytes()
    assert len(blob) > 0 and mine == blob


def test_append_and_scale(tmp_path):
    nb, bs = (3, 2, 1), (1.0, 1.0, 0.0)
    dens = np.array([0.0, 1.5, -2.0, 3.25e-7, 4.0, 0.0])
    hist = np.array([1.0, 2.0, 0.0, 4.0, 5.0, 6.0])
    out = tmp_path / "a.out"
    engine.kde_write(str(out), (0, 0, 0), bs, nb, dens, hist, time_id=1, time=2.5,
                     group_id=1, column_format=1, histogram_scale=2.0)
    engine.kde_write(str(out), (0, 0, 0), bs, nb, dens, hist, time_id=2, time=5.0,
                     group_id=1, column_format=1, histogram_scale=2.0, append=True)
    lines = out.read_text().splitlines()
    assert len(lines) == 6
    # I8 time id, ES18.9E3 time, I8 group, 3 I8 bin ids, 3 ES centres, ES
    # density, ES histogram: 8+18+8+24+54+36 = 148 columns
    assert all(len(l) == 148 for l in lines)
    assert lines[0] == ("       1  2.500000000E+000       1       2       1       1"
                        "  1.500000000E+000  5.000000000E-001  0.000000000E+000"
                        "  1.500000000E+000  4.000000000E+000")
    assert lines[3].startswith("       2  5.000000000E+000")
