"""The engine driven from Fortran on the GPU: tests/fortran/drive_engine
(built by build() from include/rwpt.f90, the bind(C) module the reference's
driver would use, and tests/fortran/drive_engine.f90) calls rwpt_create,
rwpt_set_flow, rwpt_set_dispersion, rwpt_load_particles, rwpt_advance over
12 output times, rwpt_get_particles and rwpt_kde on the config-0 problem
(100x50x1 uniform flow, 2-D walk).  Its particles must equal the oracle's
bit for bit and its density the oracle's reconstruction of those particles
within 1e-10 of the largest value."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from rwpt import problems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_engine")


@pytest.mark.skipif(not os.access(DRIVER, os.X_OK),
                    reason="Fortran driver not built (amdflang absent at build time)")
def test_engine_driven_from_fortran(tmp_path):
    P = problems.config0()
    n = 4000
    cell, lx, ly, lz = problems.release_in_cell(P, n, 1, 25, 10, seed=3, lz=0.5)
    tout = 0.25 * np.arange(1, 13)
    nb = (P.ncol, P.nrow, 1)
    origin, bins = (0.0, 0.0, 0.0), (1.0, 1.0, 0.0)
    f8, i4 = np.float64, np.int32
    hdr = np.array([P.ncol, P.nrow, P.nlay, n, tout.size, *P.dim_mask, P.advection_kind,
                    P.disp_model], i4)
    ret = np.ones(P.ncell) if P.retardation is None else P.retardation
    ib = np.ones(P.ncell, i4) if P.ibound is None else P.ibound
    icb = np.zeros(P.ncell, i4) if P.icbound is None else P.icbound
    parts = [hdr, np.array([P.ts_params[0]], f8), P.delr, P.delc, P.top, P.bot, P.frf,
             P.fff, P.flf, P.bflow, P.porosity, ret, ib, icb,
             np.array([P.default_icbound], i4), P.aL, P.aT, P.dmeff, cell, lx, ly, lz,
             tout, np.array(origin, f8), np.array(bins, f8), np.array(nb, i4)]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        for a, dt in zip(parts, [i4, f8] + [f8] * 10 + [i4, i4, i4, f8, f8, f8, i4] +
                         [f8] * 4 + [f8, f8, i4]):
            fh.write(np.ascontiguousarray(a, dtype=dt).tobytes())
    print("fortran driver: config0, 4000 particles, 12 intervals", flush=True)
    subprocess.run([DRIVER, str(fin), str(fout)], check=True, timeout=120)
    raw = open(fout, "rb").read()
    off = 0

    def take(dt, k):
        nonlocal off
        a = np.frombuffer(raw[off:off + k * np.dtype(dt).itemsize], dt).copy()
        off += k * np.dtype(dt).itemsize
        return a

    rc = take(i4, 1)[0]
    assert rc == 0
    got = {"cell": take(i4, n), "status": take(i4, n)}
    for f in ("lx", "ly", "lz", "t"):
        got[f] = take(f8, n)
    ncells = int(np.prod(nb))
    dens, hist = take(f8, ncells), take(f8, ncells)
    st = O.new_state(cell, lx, ly, lz)
    OP = O.OracleProblem(P)
    t0 = 0.0
    for t1 in tout:
        assert O.advance(OP, st, t0, t1) == 0
        t0 = t1
    for f in ("cell", "status", "lx", "ly", "lz", "t"):
        assert np.array_equal(got[f], st[f]), (f, int((got[f] != st[f]).sum()))
    on = st["status"] == 1
    gx, gy, gz = O.global_xyz(P, st["cell"][on], st["lx"][on], st["ly"][on], st["lz"][on])
    wd, wc, _ = O.kde_density(gx, gy, gz, origin, bins, nb)
    assert np.array_equal(hist, wc)
    assert np.abs(dens - wd).max() <= 1e-10 * wd.max()
    assert on.sum() > n // 2 and hist.sum() == on.sum()
