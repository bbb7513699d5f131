"""FP32-local walk (precision 1 and 2) against the FP64 walk.

precision 1 consumes the FP64 walk's random draws (12 per normal, summed
exactly in integers, rounded once to float), so an FP32 particle follows the
FP64 particle's path until a rounding difference flips a branch: most
particles end in the same cell with the same status, and the plume moments
differ by the FP32 arithmetic alone (bounded here at 0.5 %).

precision 2 builds each normal from 4 draws, so its particles follow other
paths: its moments are compared with the FP64 moments through their
sampling errors (a two-sample test at 4.5 standard errors).
"""
import numpy as np
import pytest

from oracle import oracle as O
from rwpt import Engine, problems
from scenarios import SCEN, release

pytestmark = pytest.mark.gpu


def _run(P, rel, precision, times, stream_offset=0):
    E = Engine(P, precision=precision)
    out = []
    try:
        E.load_particles(*rel, stream_offset=stream_offset)
        for t in times:
            E.advance(t)
            out.append(E.particles())
    finally:
        E.close()
    return out


def _coords(P, g):
    act = g["status"] == 1
    return np.stack(O.global_xyz(P, g["cell"][act], g["lx"][act], g["ly"][act],
                                 g["lz"][act]))


@pytest.mark.parametrize("name", ["hetero3d", "hetero_ibound_rebound_R", "cfg0",
                                  "axisymmetric", "backward", "min_adv_disp"])
def test_fp32_same_draws_follow_fp64_paths(name):
    P, where, dt_out, nout = SCEN[name]
    rel = release(P, where, 20000)
    times = [dt_out * k for k in range(1, nout + 1)]
    a = _run(P, rel, 0, times)
    b = _run(P, rel, 1, times)
    g0, g1 = a[-1], b[-1]
    same = (g0["cell"] == g1["cell"]) & (g0["status"] == g1["status"])
    assert same.mean() >= 0.97
    # the state domains hold in FP32 too
    on = g1["status"] == 1
    for f in ("lx", "ly", "lz"):
        assert np.all((g1[f][on] >= 0.0) & (g1[f][on] <= 1.0))
    assert np.all(g1["t"][on] == times[-1])
    # where the paths agree, local positions agree to FP32 rounding growth
    both = same & on
    d = np.abs(g1["lx"][both] - g0["lx"][both])
    assert np.median(d) < 1e-4
    # statuses agree in number
    c0 = np.bincount(g0["status"], minlength=10)
    c1 = np.bincount(g1["status"], minlength=10)
    assert np.abs(c1 - c0).max() <= max(5, 0.005 * len(rel[0]))


def test_fp32_moments_config1():
    """Heterogeneous 200x200x20 field, 1e6 particles, 2 tracking intervals
    (100 RW steps): relative errors of the plume mean and variance."""
    P = problems.config1()
    rel = problems.release_in_block(P, 1_000_000, (1, 10), (76, 125), (1, 20))
    times = [25.0, 50.0]
    r0 = _run(P, rel, 0, times)
    r1 = _run(P, rel, 1, times)
    r2 = _run(P, rel, 2, times)
    for k in range(len(times)):
        x0, x1, x2 = (_coords(P, r[k]) for r in (r0, r1, r2))
        m0, v0 = x0.mean(axis=1), x0.var(axis=1)
        m1, v1 = x1.mean(axis=1), x1.var(axis=1)
        assert np.all(np.abs(m1 - m0) <= 5e-3 * np.abs(m0))
        assert np.all(np.abs(v1 - v0) <= 5e-3 * v0)
        # two-sample test of precision 2 on means and variances
        n0, n2 = x0.shape[1], x2.shape[1]
        m2, v2 = x2.mean(axis=1), x2.var(axis=1)
        se_m = np.sqrt(v0 / n0 + v2 / n2)
        k0 = ((x0 - m0[:, None]) ** 4).mean(axis=1)
        k2 = ((x2 - m2[:, None]) ** 4).mean(axis=1)
        se_v = np.sqrt((k0 - v0 ** 2) / n0 + (k2 - v2 ** 2) / n2)
        assert np.all(np.abs(m2 - m0) <= 4.5 * se_m), (m2 - m0) / se_m
        assert np.all(np.abs(v2 - v0) <= 4.5 * se_v), (v2 - v0) / se_v


@pytest.mark.parametrize("precision", [1, 2])
def test_fp32_exponential_advection(precision):
    """EXPONENTIAL advection with the FP64 Newton exit solve: plume means
    within sampling error of the FP64 walk."""
    P, where, dt_out, nout = SCEN["hetero3d"]
    P = problems.Problem(**{**P.__dict__, "advection_kind": 1})
    rel = release(P, where, 20000)
    times = [dt_out * k for k in range(1, nout + 1)]
    g0 = _run(P, rel, 0, times)[-1]
    g1 = _run(P, rel, precision, times)[-1]
    x0, x1 = _coords(P, g0), _coords(P, g1)
    se = np.sqrt(x0.var(axis=1) / x0.shape[1] + x1.var(axis=1) / x1.shape[1])
    assert np.all(np.abs(x1.mean(axis=1) - x0.mean(axis=1)) <= 4.5 * se)
    if precision == 1:
        same = (g0["cell"] == g1["cell"]) & (g0["status"] == g1["status"])
        assert same.mean() >= 0.9


def test_fp32_launch_shape_does_not_change_results():
    """The resumable FP32 walk is deterministic under any budget."""
    P, where, dt_out, nout = SCEN["hetero_ibound_rebound_R"]
    rel = release(P, where, 3000)
    out = []
    for budget in (1, 64):
        E = Engine(P, precision=2)
        try:
            E.set_tuning(budget, 128, 1, 1)
            E.load_particles(*rel)
            for k in range(1, 4):
                E.advance(k * dt_out)
            out.append(E.particles())
        finally:
            E.close()
    for f in ("cell", "status", "lx", "ly", "lz", "t", "rng"):
        assert np.array_equal(out[0][f], out[1][f]), f
