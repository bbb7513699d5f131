"""The benchmark configurations at their real sizes on the device.

* the 200x200x20 heterogeneous field with 1e6 particles: one tracking
  interval (50 RW steps) of all particles on the device, and the oracle on a
  20,000-particle sample of them with the same rng streams, bit for bit;
* the same field with 1e7 particles, two intervals: the FP32-local walk
  against the plain FP64 walk on the plume mean and variance (<= 0.5 %
  relative), state domains and a repeat of the FP32 run bit for bit;
* 256 ensemble members of 100x100x10 (about 6.3e7 particles), one interval
  in FP32: state domains, and three sampled members equal to the same
  members run alone bit for bit;
* the calibration workload (8 zones on the 200x200x20 field, 1e6
  particles): the adjoint's forward pass leaves the particles exactly where
  rwpt_advance does, and its 4-step gradient matches central differences.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rwpt import Engine, problems

pytestmark = pytest.mark.gpu

FIELDS = ("cell", "status", "lx", "ly", "lz", "t", "substeps", "rng")
INTERVAL = 25.0   # 50 RW steps of 0.5 d


def _domains(g, ncell):
    on = g["status"] == 1
    for f in ("lx", "ly", "lz"):
        assert np.all((g[f] >= 0.0) & (g[f] <= 1.0)), f
    assert np.all((g["cell"] >= 1) & (g["cell"] <= ncell))
    assert set(np.unique(g["status"])) <= {1, 2, 9}
    assert (g["status"] == 9).mean() < 1e-4
    return on


def test_config1_sample_bitexact():
    P = problems.config1()
    n = 1_000_000
    rel = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    print("config1: 1e6 particles, one interval; oracle on every 50th", flush=True)
    E = Engine(P)
    try:
        E.load_particles(*rel)
        E.advance(INTERVAL)
        g = E.particles()
    finally:
        E.close()
    idx = np.arange(0, n, 50)
    streams = O.rng_streams(n).reshape(n, 2)[idx].ravel()
    st = O.new_state(*(a[idx] for a in rel), rng=streams)
    assert O.advance(O.OracleProblem(P), st, 0.0, INTERVAL, nthreads=0) == 0
    for f in FIELDS:
        got = g[f].reshape(n, 2)[idx].ravel() if f == "rng" else g[f][idx]
        assert np.array_equal(got, st[f]), (f, int((got != st[f]).sum()))
    assert st["substeps"].sum() > 20 * idx.size
    assert (st["status"] == 1).mean() > 0.5


def test_config2_fp32_against_fp64():
    P = problems.config1()
    n = 10_000_000
    rel = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    print("config2: 1e7 particles, two intervals, FP32 and FP64", flush=True)
    out = {}
    for prec, reps in ((1, 2), (0, 1)):
        E = Engine(P, precision=prec)
        try:
            runs = []
            for _ in range(reps):
                E.load_particles(*rel)
                E.set_time(0.0)
                E.advance(INTERVAL)
                E.advance(2 * INTERVAL)
                g = E.particles()
                gx, gy, gz = E.global_xyz()
                runs.append((g, np.stack([gx, gy, gz])))
        finally:
            E.close()
        if reps == 2:
            for f in FIELDS:
                assert np.array_equal(runs[0][0][f], runs[1][0][f]), ("repeat", f)
        out[prec] = runs[0]
    moments = {}
    for prec, (g, x) in out.items():
        on = _domains(g, P.ncell)
        assert on.mean() > 0.5
        moments[prec] = (x[:, on].mean(axis=1), x[:, on].var(axis=1))
    mean_err = np.abs(moments[1][0] - moments[0][0]) / np.abs(moments[0][0])
    var_err = np.abs(moments[1][1] - moments[0][1]) / moments[0][1]
    print("config2 errors: mean", mean_err, "variance", var_err, flush=True)
    assert mean_err.max() <= 0.005 and var_err.max() <= 0.005


def test_config3_full_ensemble():
    E, counts = problems.config3()
    rel = problems.release_ensemble(E, counts, (1, 5), (38, 63), (1, 10))
    total = int(counts.sum())
    print(f"config3: {E.members} members, {total} particles, FP32", flush=True)
    eng = Engine(E, precision=1)
    try:
        eng.load_particles(*rel)
        act, steps = eng.advance(INTERVAL)
        g = eng.particles()
    finally:
        eng.close()
    on = _domains(g, E.ncell)
    assert act == int(on.sum()) and steps > 10 * total
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    for m in (0, E.members // 2, E.members - 1):
        P = problems.member(E, m)
        s = slice(int(offs[m]), int(offs[m] + counts[m]))
        cell = rel[0][s] - m * E.ncell_member
        one = Engine(P, precision=1)
        try:
            one.load_particles(cell, rel[1][s], rel[2][s], rel[3][s],
                               stream_offset=int(offs[m]))
            one.advance(INTERVAL)
            w = one.particles()
        finally:
            one.close()
        print(f"  member {m}: {counts[m]} particles", flush=True)
        assert np.array_equal(g["cell"][s] - m * E.ncell_member, w["cell"]), m
        for f in ("status", "lx", "ly", "lz", "t", "substeps"):
            assert np.array_equal(g[f][s], w[f]), (m, f)
        assert np.array_equal(g["rng"][2 * s.start:2 * s.stop], w["rng"]), m


def test_config4_adjoint_forward_and_short_gradient():
    P, zones, (aL, aT, _, dm) = problems.config4(ts_params=(0.25, 0.0))
    n = 1_000_000
    rel = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    windows = [0.25 * 25 * (k + 1) for k in range(2)]
    obs = problems.observation_points(P)
    print("config4: 1e6 particles, forward of 2 windows, 4-step gradient", flush=True)

    def engine(a_l):
        return Engine(problems.zoned_dispersion(P, zones, a_l, aT, dm))

    E = engine(aL)
    try:
        E.load_particles(*rel)
        E.adjoint(windows, obs, np.zeros(len(obs)), (2.0, 2.0, 1.0), gradient=False)
        a = E.particles()
    finally:
        E.close()
    E = engine(aL)
    try:
        E.load_particles(*rel)
        for t in windows:
            E.advance(t)
        b = E.particles()
    finally:
        E.close()
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    # the gradient over 4 RW steps against central differences with common
    # random numbers (no particle changes branch at this step)
    short = [1.0]

    def misfit(a_l, grad=False, target=None):
        Q = engine(a_l)
        try:
            Q.load_particles(*rel)
            return Q.adjoint(short, obs, np.zeros(len(obs)) if target is None else target,
                             (2.0, 2.0, 1.0), gradient=grad)
        finally:
            Q.close()

    ref = misfit(aL)
    target = 0.8 * ref["conc"]
    A = misfit(aL, grad=True, target=target)
    g = float(A["grad_al"][zones == 0].sum())
    h = 1e-8 * aL[0]
    vals = []
    for sg in (1, -1):
        a2 = aL.copy()
        a2[0] += sg * h
        vals.append(misfit(a2, target=target)["misfit"])
    fd = (vals[0] - vals[1]) / (2 * h)
    print(f"config4 short gradient {g:.10g} central difference {fd:.10g}", flush=True)
    assert g != 0.0 and abs(g - fd) <= 1e-4 * abs(fd)
