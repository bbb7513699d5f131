"""Pins of the oracle's random-walk step against closed forms and brute force.

* pure advection in uniform flow is exact (x = x0 + v t);
* the 2-D and 3-D Gaussian plume in uniform flow: mean x0 + v t, variances
  2(alpha v + Dm) t / R (run at a time step where the jump limiter of
  TrackSubCell.F90:989-1013 does not fire, see test_jump_limiter_bias);
* EXPONENTIAL and EULERIAN advection coincide bit for bit when the velocity
  gradient is below 1e-10 (TrackSubCell.F90:1487-1538);
* corner discharge, corner porosity and corner dispersion recomputed from
  the grid geometry (faces touching each corner point) rather than from the
  reference's slot tables, with the two reference traps applied explicitly:
  corners 110/111 swap their porosity stencils (TrackSubCell.F90:4329-4363)
  and corners 010..111 use |q| of corner 100 in alphaT*|q|
  (TrackSubCell.F90:2622-2677).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from rwpt import problems
from scenarios import SCEN, release


def _run(P, n, t_end, steps=1, seed=3, lz=0.5, cellkij=(1, 25, 10), rng=None):
    cell, lx, ly, lz_ = problems.release_in_cell(P, n, *cellkij, seed=seed, lz=lz)
    st = O.new_state(cell, lx, ly, lz_, rng=rng)
    OP = O.OracleProblem(P)
    t = 0.0
    for s in range(1, steps + 1):
        t1 = t_end * s / steps
        assert O.advance(OP, st, t, t1) == 0
        t = t1
    return st


def test_pure_advection_exact():
    P = problems.config0(aL=np.array([0.0]), aT=np.array([0.0]),
                         dmeff=np.array([0.0]))
    st = _run(P, 200, 5.0, steps=20)
    gx, gy, _ = O.global_xyz(P, st["cell"], st["lx"], st["ly"], st["lz"])
    cell0, lx0, ly0, _ = problems.release_in_cell(P, 200, 1, 25, 10, seed=3, lz=0.5)
    x0 = 9.0 + lx0
    assert np.all(st["status"] == 1)
    np.testing.assert_allclose(gx, x0 + 5.0 / 0.3, rtol=0, atol=1e-11)
    # row 25 of 50 spans y in [25, 26] (Y(1) is the back edge at y = 50)
    np.testing.assert_allclose(gy, 25.0 + ly0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("kind", [1, 2])
def test_plume_2d_moments(kind):
    # config-0 field at dt = 0.01 d (jump <= 0.5 cell), t = 10 d
    P = problems.config0(ts_params=(0.01, 0.0), advection_kind=kind)
    st = _run(P, 6000, 10.0)
    gx, gy, _ = O.global_xyz(P, st["cell"], st["lx"], st["ly"], st["lz"])
    v, t = 1.0 / 0.3, 10.0
    assert np.all(st["status"] == 1)
    assert abs(gx.mean() - (9.5 + v * t)) / (v * t) < 0.02
    assert abs(gy.mean() - 25.5) < 0.05
    vx_exp = 1 / 12 + 2 * (0.1 * v + 1e-4) * t
    vy_exp = 1 / 12 + 2 * (0.01 * v + 1e-4) * t
    assert abs(gx.var() / vx_exp - 1) < 0.06
    assert abs(gy.var() / vy_exp - 1) < 0.07


def test_plume_3d_moments():
    P = problems.uniform_flow(60, 21, 21, qx=1.0, porosity=0.25)
    P.aL = np.array([0.2])
    P.aT = np.array([0.05])
    P.dmeff = np.array([1e-3])
    P.ts_params = (0.01, 0.0)
    st = _run(P, 4000, 6.0, cellkij=(11, 11, 5), lz=None)
    gx, gy, gz = O.global_xyz(P, st["cell"], st["lx"], st["ly"], st["lz"])
    v, t = 4.0, 6.0
    assert np.all(st["status"] == 1)
    assert abs(gx.mean() - (4.5 + v * t)) / (v * t) < 0.02
    for g, a in ((gx, 0.2), (gy, 0.05), (gz, 0.05)):
        exp = 1 / 12 + 2 * (a * v + 1e-3) * t
        assert abs(g.var() / exp - 1) < 0.08


def test_exponential_equals_eulerian_without_exits():
    # 100 m cells: no particle reaches a face, so only the advection
    # displacement differs between the two modes -- and it must not
    def prob(kind):
        P = problems.uniform_flow(3, 3, 1, qx=1.0, porosity=0.3, d=100.0)
        P.aL, P.aT, P.dmeff = np.array([0.1]), np.array([0.01]), np.array([1e-4])
        P.dim_mask, P.ts_params, P.advection_kind = (1, 1, 0), (0.25, 0.0), kind
        return P

    def run(P):
        rng = np.random.default_rng(9)
        n = 300
        st = O.new_state(np.full(n, 5, np.int32), rng.uniform(0.3, 0.6, n),
                         rng.uniform(0.3, 0.7, n), np.full(n, 0.5))
        for s in range(1, 9):
            O.advance(P, st, (s - 1) * 0.25, s * 0.25)
        return st
    a = run(prob(1))
    b = run(prob(2))
    assert np.all(a["cell"] == 5)
    for k in ("cell", "status", "lx", "ly", "lz", "t", "substeps", "rng"):
        assert np.array_equal(a[k], b[k]), k


def _f_exp(dt, v, g, dI, divD, dB):
    dA = v * (np.exp(g * dt) - 1) / g if abs(g) > 1e-10 else v * dt
    return dA + divD * dt + dB * np.sqrt(dt) - dI


@pytest.mark.parametrize("seed", range(5))
def test_newton_exit_time_is_a_root(seed):
    """NewtonRaphsonTimeStepExponentialAdvection returns a root of
    v(e^{g dt}-1)/g + divD dt + dB sqrt(dt) = dI; with g = 0 it equals the
    closed-form quadratic root of the EULERIAN solver."""
    L = O.lib()
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(300):
        x = rng.uniform(0.5, 0.99)
        v1, v2 = rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0)
        v = (1 - x) * v1 + x * v2
        divD = rng.normal(0, 0.05)
        dB = rng.normal(0, 0.5)
        dt = 0.3
        dI = 1.0 * (1 - x)
        if _f_exp(dt, v, v2 - v1, dI, divD, dB) <= 0:  # must overshoot
            continue
        r = L.orc_newton_axis(dt, v, v1, v2, 1.0, dI, divD, dB)
        if r == 0.0 or r == dt:
            continue
        assert abs(_f_exp(r, v, v2 - v1, dI, divD, dB)) <= 1e-7 * dI
        # uniform velocity: Newton vs the quadratic in sqrt(dt)
        ef = C.c_int(0)
        q = L.orc_quadratic_axis(dt, 1.5, x, 1.0, v, divD, dB, C.byref(ef))
        rn = L.orc_newton_axis(dt, v, v, v, 1.0, dI, divD, dB)
        if q > 0 and 0 < rn < dt:
            assert abs(rn - q) <= 1e-6 * q
        checked += 1
    assert checked > 100


def test_jump_limiter_bias():
    """At config 0's FIXED dt = 0.25 d (Courant 0.83) the limiter reuses the
    draw that failed to pick dt, which slows the plume (reference behaviour
    kept as is).  The test documents the size of the effect."""
    P = problems.config0()
    st = _run(P, 3000, 5.0, steps=20)
    gx, _, _ = O.global_xyz(P, st["cell"], st["lx"], st["ly"], st["lz"])
    ratio = (gx.mean() - 9.5) / (5.0 / 0.3)
    assert 0.55 < ratio < 0.8


def test_open_boundary_exit():
    P = problems.config0(ts_params=(0.05, 0.0))
    st = _run(P, 300, 40.0, steps=4, cellkij=(1, 25, 90))
    out = st["status"] == 2
    assert out.sum() > 250
    assert np.all(st["lx"][out] == 1.0)
    cols = (st["cell"][out] - 1) % 100 + 1
    assert np.all(cols == 100)


def test_reflecting_boundaries_keep_particles():
    P = problems.config0(ts_params=(0.05, 0.0))
    P.default_icbound = 1
    P.aT = np.array([0.5])
    st = _run(P, 300, 30.0, steps=3, cellkij=(1, 2, 95))
    assert np.all(st["status"] == 1)
    assert np.all((st["lx"] >= 0) & (st["lx"] <= 1) & (st["ly"] >= 0) & (st["ly"] <= 1))


# --------------------------------------------------------------------------
# corner tables by geometry
# --------------------------------------------------------------------------

def _assembled_faces(P):
    """Q1..Q6 per cell from MODFLOW arrays + boundary flows (no inactive)."""
    nl, nr, nc = P.nlay, P.nrow, P.ncol
    frf = P.frf.reshape(nl, nr, nc)
    fff = P.fff.reshape(nl, nr, nc)
    flf = P.flf.reshape(nl, nr, nc)
    Q = np.zeros((nl, nr, nc, 6))
    Q[:, :, 1:, 0] = frf[:, :, :-1]
    Q[:, :, :, 1] = frf
    Q[:, :, :, 2] = -fff
    Q[:, 1:, :, 3] = -fff[:, :-1, :]
    Q[:, :, :, 4] = -flf
    Q[1:, :, :, 5] = -flf[:-1, :, :]
    if P.bflow is not None:
        bf = P.bflow.reshape(nl, nr, nc, 6)
        sign = np.array([1, -1, 1, -1, 1, -1])
        edge = [(slice(None), slice(None), 0), (slice(None), slice(None), -1),
                (slice(None), -1, slice(None)), (slice(None), 0, slice(None)),
                (-1, slice(None), slice(None)), (0, slice(None), slice(None))]
        for f in range(6):
            m = bf[..., f] != 0
            e = np.zeros_like(m)
            e[edge[f]] = True
            Q[..., f] = np.where(m & e, sign[f] * bf[..., f], Q[..., f])
    return Q


def _geometric_tables(P):
    nl, nr, nc = P.nlay, P.nrow, P.ncol
    Q = _assembled_faces(P)
    dz = (P.top - P.bot).reshape(nl, nr, nc)
    dx = np.broadcast_to(P.delr[None, None, :], (nl, nr, nc))
    dy = np.broadcast_to(P.delc[None, :, None], (nl, nr, nc))
    phi = P.porosity.reshape(nl, nr, nc)
    aL, aT, dm = P.aL[0], P.aT[0], P.dmeff[0]
    out = np.zeros((nl, nr, nc, 8, 4))
    pc = np.zeros((nl, nr, nc, 8))
    D = np.zeros((nl, nr, nc, 6, 8))
    for k in range(nl):
        for i in range(nr):
            for j in range(nc):
                for ci in range(8):
                    cx, cy, cz = ci & 1, (ci >> 1) & 1, (ci >> 2) & 1
                    rows = [i, i + 1] if cy == 0 else [i, i - 1]    # y- is row+1
                    lays = [k, k + 1] if cz == 0 else [k, k - 1]    # z- is layer+1
                    cols = [j, j - 1] if cx == 0 else [j, j + 1]
                    def ok(kk, ii, jj):
                        return 0 <= kk < nl and 0 <= ii < nr and 0 <= jj < nc
                    # x component: x-faces of this column touching the corner
                    fsum = asum = 0.0
                    for ii in rows:
                        for kk in lays:
                            if ok(kk, ii, j):
                                fsum += Q[kk, ii, j, cx]
                                asum += dy[kk, ii, j] * dz[kk, ii, j]
                    out[k, i, j, ci, 0] = fsum / asum
                    fsum = asum = 0.0
                    for jj in cols:
                        for kk in lays:
                            if ok(kk, i, jj):
                                fsum += Q[kk, i, jj, 2 + cy]
                                asum += dx[kk, i, jj] * dz[kk, i, jj]
                    out[k, i, j, ci, 1] = fsum / asum
                    fsum = asum = 0.0
                    for jj in cols:
                        for ii in rows:
                            if ok(k, ii, jj):
                                fsum += Q[k, ii, jj, 4 + cz]
                                asum += dx[k, ii, jj] * dy[k, ii, jj]
                    out[k, i, j, ci, 2] = fsum / asum
                    # porosity: center + stencil cells touching the corner
                    # (the one diagonal cell is outside the 18-cell stencil)
                    vs = dx[k, i, j] * dy[k, i, j] * dz[k, i, j]
                    ws = phi[k, i, j] * vs
                    for kk in lays:
                        for ii in rows:
                            for jj in cols:
                                if (kk, ii, jj) == (k, i, j):
                                    continue
                                if kk != k and ii != i and jj != j:
                                    continue
                                if ok(kk, ii, jj):
                                    v = dx[kk, ii, jj] * dy[kk, ii, jj] * dz[kk, ii, jj]
                                    vs += v
                                    ws += v * phi[kk, ii, jj]
                    pc[k, i, j, ci] = ws / vs
                out[k, i, j, :, 3] = np.sqrt((out[k, i, j, :, :3] ** 2).sum(-1))
                # reference trap: corners 110 and 111 swap porosity stencils
                pc[k, i, j, [3, 7]] = pc[k, i, j, [7, 3]]
                if P.uniform_porosity:
                    pc[k, i, j, :] = phi[k, i, j]
                q = out[k, i, j]
                for ci in range(8):
                    qn = q[ci, 3]
                    qn_t = q[0, 3] if ci == 0 else q[1, 3]  # trap T1
                    base = aT * qn_t + pc[k, i, j, ci] * dm
                    p = np.zeros(6) if qn == 0 else np.array(
                        [q[ci, 0] ** 2, q[ci, 1] ** 2, q[ci, 2] ** 2,
                         q[ci, 0] * q[ci, 1], q[ci, 0] * q[ci, 2],
                         q[ci, 1] * q[ci, 2]]) / qn
                    D[k, i, j, :, ci] = [base + (aL - aT) * p[0], base + (aL - aT) * p[1],
                                         base + (aL - aT) * p[2], (aL - aT) * p[3],
                                         (aL - aT) * p[4], (aL - aT) * p[5]]
    return out.reshape(-1, 8, 4), pc.reshape(-1, 8), D.reshape(-1, 6, 8)


@pytest.mark.parametrize("shape", [(2, 2, 1), (3, 3, 3), (5, 4, 3)])
def test_corner_tables_match_geometry(shape):
    nc, nr, nl = shape
    P = problems.heterogeneous_flow(nc, nr, nl, corr_len=2.0, seed=11)
    P.aL, P.aT, P.dmeff = np.array([0.3]), np.array([0.04]), np.array([2e-3])
    tab = O.cell_tables(P)
    q, pc, D = _geometric_tables(P)
    np.testing.assert_allclose(tab[:, 12:44].reshape(-1, 8, 4), q, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(tab[:, 44:52], pc, rtol=1e-13)
    np.testing.assert_allclose(tab[:, 52:100].reshape(-1, 6, 8), D, rtol=1e-12, atol=1e-15)
    # face pore velocities V = Q / (area phi R)
    Qa = _assembled_faces(P).reshape(-1, 6)
    np.testing.assert_allclose(tab[:, 0], Qa[:, 0] / P.porosity, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(tab[:, 5], Qa[:, 5] / P.porosity, rtol=1e-14, atol=1e-16)


def test_heterogeneous_field_is_divergence_free():
    P = problems.heterogeneous_flow(20, 16, 6, corr_len=3.0, seed=5)
    Q = _assembled_faces(P)
    div = Q[..., 1] - Q[..., 0] + Q[..., 3] - Q[..., 2] + Q[..., 5] - Q[..., 4]
    assert np.abs(div).max() < 1e-12


@pytest.mark.parametrize("name", ["stop_zone", "weak_sinks", "weak_sources_backward",
                                  "transient_weak_sinks", "transient_no_exit",
                                  "transient_weak_sources_backward"])
def test_stopping_scenarios_reach_their_statuses(name):
    """The stopping scenarios stop particles with the status they exercise
    (checked on the oracle, which the bit-exact tests tie to the device)."""
    P, where, dt_out, nout = SCEN[name]
    cell, lx, ly, lz = release(P, where, 1500)
    st = O.new_state(cell, lx, ly, lz)
    OP = O.OracleProblem(P)
    t = 0.0
    for k in range(1, nout + 1):
        O.advance(OP, st, t, k * dt_out)
        t = k * dt_out
    want = {"stop_zone": 6, "weak_sinks": 3, "weak_sources_backward": 4,
            "transient_weak_sinks": 3, "transient_no_exit": 5,
            "transient_weak_sources_backward": 4}[name]
    assert (st["status"] == want).sum() > 50
