"""Pathwise adjoint of the observation misfit (rwpt_adjoint).

The forward part must advance exactly as rwpt_advance does (bit for bit),
and the gradient must be the derivative of the misfit with common random
numbers: for a step small enough that no particle changes branch, the
central finite difference of the misfit over rebuilt engines equals the
pathwise derivative up to O(h^2) and rounding; particles near a grazing
exit have large second derivatives, so the two are required to agree to
1e-4 relative at one of the steps tried.  A branch change (a jump-limiter cut or a cell exit taken on one
side only) moves a particle by a fraction of a cell instead of O(h), so the
check first compares the two runs' final cells, statuses, sub-step counts
and positions.  Parameters: per-zone aL and aT (zones = column bands), a
common Dm and a common porosity scale."""
import numpy as np
import pytest

from rwpt import Engine, RwptError, problems

pytestmark = pytest.mark.gpu

WINDOWS = (1.0, 2.0, 3.0)
NZ = 3


def base(ndim=3):
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.ts_params = (0.5, 0.0)
    P.time_step_kind = 4
    if ndim == 2:
        P.dim_mask = (1, 1, 0)
    zones = problems.column_zones(P, NZ)
    aL = np.array([0.3, 0.5, 0.4])
    aT = np.array([0.04, 0.05, 0.03])
    return P, zones, aL, aT, 1e-3


def release(P):
    return problems.release_in_block(P, 400, (2, 6), (6, 15), (1, 6), seed=5)


def observations(P):
    return problems.observation_points(P, 6, 3, (5, 18), (7, 14))


def run(P, zones, aL, aT, dm, por_scale=1.0, gradient=False, target=None):
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    Q.porosity = P.porosity * por_scale
    E = Engine(Q)
    try:
        E.load_particles(*release(P))
        obs = observations(P)
        tgt = np.zeros(len(obs)) if target is None else target
        out = E.adjoint(WINDOWS, obs, tgt, (1.5, 1.5, 1.0), gradient=gradient)
        g = E.particles()
        out["state"] = (g["cell"].copy(), g["status"].copy(), g["substeps"].copy())
        out["pos"] = np.stack([g["lx"], g["ly"], g["lz"], g["t"]])
    finally:
        E.close()
    return out


def same_branches(a, b, rel):
    same = all(np.array_equal(x, y) for x, y in zip(a["state"], b["state"]))
    return same and float(np.max(np.abs(a["pos"] - b["pos"]))) < 1e4 * rel


def central_difference(P, zones, aL, aT, dm, target, name, z, rel):
    """(J(p + h) - J(p - h)) / 2h for one parameter and h = rel * p, or None
    when a particle changed branch between the two runs."""
    vals = []
    for sgn in (1.0, -1.0):
        aL2, aT2, dm2, s2 = aL.copy(), aT.copy(), dm, 1.0
        if name == "aL":
            h = rel * aL[z]
            aL2[z] += sgn * h
        elif name == "aT":
            h = rel * aT[z]
            aT2[z] += sgn * h
        elif name == "dm":
            h = rel * dm
            dm2 = dm + sgn * h
        else:
            h = rel
            s2 = 1.0 + sgn * h
        vals.append(run(P, zones, aL2, aT2, dm2, s2, target=target))
    if not same_branches(vals[0], vals[1], rel):
        return None
    return (vals[0]["misfit"] - vals[1]["misfit"]) / (2 * h), h


@pytest.mark.parametrize("ndim", [3, 2])
def test_adjoint_matches_central_differences(ndim):
    P, zones, aL, aT, dm = base(ndim)
    print(f"adjoint ndim {ndim}: 400 particles, windows {WINDOWS}", flush=True)
    ref = run(P, zones, aL, aT, dm)
    target = 0.7 * ref["conc"][::-1]
    A = run(P, zones, aL, aT, dm, gradient=True, target=target)
    assert A["misfit"] > 0
    for k in ("grad_al", "grad_at", "grad_dm"):
        assert np.isfinite(A[k]).all(), (k, int((~np.isfinite(A[k])).sum()))
    assert np.isfinite(A["grad_porosity_scale"])
    adj = {}
    for z in range(NZ):
        adj[("aL", z)] = A["grad_al"][zones == z].sum()
        adj[("aT", z)] = A["grad_at"][zones == z].sum()
    adj[("dm", 0)] = A["grad_dm"].sum()
    adj[("por", 0)] = A["grad_porosity_scale"]
    checked = 0
    for (name, z), g in adj.items():
        # steps without a branch change; a particle near a grazing exit has
        # a large second derivative, so one of the steps may still be outside
        # the linear range: the adjoint must agree with at least one of them,
        # allowing the misfit's rounding (order of the atomic sums, ~1e-15
        # relative) divided by 2h
        seen = []
        for rel in (1e-6, 1e-7, 1e-8, 1e-9):
            r = central_difference(P, zones, aL, aT, dm, target, name, z, rel)
            if r is None:
                print(f"  {name}{z} rel {rel:g}: a particle changed branch", flush=True)
                continue
            fd, h = r
            print(f"  {name}{z} rel {rel:g}: adjoint {g:.9e} fd {fd:.9e}", flush=True)
            seen.append(fd)
            if abs(g - fd) <= 1e-4 * abs(fd) + 1e-14 * A["misfit"] / h:
                checked += 1
                break
        else:
            assert not seen, (name, z, g, seen)
    assert checked >= len(adj) - 1


def test_forward_part_is_the_advance():
    """The misfit call advances bit-identically to rwpt_advance, and the
    gradient leaves the same final state."""
    P, zones, aL, aT, dm = base()
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    E1, E2 = Engine(Q), Engine(Q)
    try:
        for E in (E1, E2):
            E.load_particles(*release(P))
        obs = observations(P)
        out = E1.adjoint(WINDOWS, obs, np.zeros(len(obs)), (1.5, 1.5, 1.0), gradient=True)
        for t in WINDOWS:
            E2.advance(t)
        a, b = E1.particles(), E2.particles()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
        gx, gy, gz = E2.global_xyz()
        on = b["status"] == 1
        c = np.zeros(len(obs))
        for o, (x, y, z) in enumerate(obs):
            k = (np.exp(-0.5 * ((gx[on] - x) / 1.5) ** 2) / (1.5 * np.sqrt(2 * np.pi)) *
                 np.exp(-0.5 * ((gy[on] - y) / 1.5) ** 2) / (1.5 * np.sqrt(2 * np.pi)) *
                 np.exp(-0.5 * ((gz[on] - z) / 1.0) ** 2) / (1.0 * np.sqrt(2 * np.pi)))
            c[o] = k.sum()
        np.testing.assert_allclose(out["conc"], c, rtol=1e-12, atol=1e-300)
        assert out["misfit"] == pytest.approx((c ** 2).sum(), rel=1e-12)
    finally:
        E1.close()
        E2.close()


def test_adjoint_rejects_unsupported():
    P, zones, aL, aT, dm = base()
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    E = Engine(Q, precision=1)
    try:
        E.load_particles(*release(P))
        obs = observations(P)
        with pytest.raises(RwptError):
            E.adjoint(WINDOWS, obs, np.zeros(len(obs)), (1.5, 1.5, 1.0), gradient=True)
    finally:
        E.close()


def test_autograd_wrapper():
    """ZonedMisfit: the forward is the engine's misfit, the backward its
    zone-summed adjoint gradient (porosity scale and Dm included)."""
    import torch
    from rwpt.adjoint import CalibrationProblem, ZonedMisfit
    P, zones, aL, aT, dm = base()
    obs = observations(P)
    prob = CalibrationProblem(P, zones, release(P), WINDOWS, obs, np.zeros(len(obs)),
                              (1.5, 1.5, 1.0))
    theta0 = np.concatenate([aL, aT, [1.0, dm]])
    ref, _ = prob.evaluate(theta0, gradient=False)
    prob.target = 0.7 * np.sqrt(ref / len(obs)) * np.ones(len(obs))
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    J = ZonedMisfit.apply(theta, prob)
    J.backward()
    Jn, g = prob.evaluate(theta0, gradient=True)
    # per-cell gradients and concentrations are sums of FP64 atomics, so two
    # evaluations agree to rounding, not bit for bit
    assert float(J) == pytest.approx(Jn, rel=1e-12)
    np.testing.assert_allclose(theta.grad.numpy(), g, rtol=1e-9, atol=0)
    assert np.all(np.isfinite(g)) and np.any(g != 0)
    # a change of the porosity scale: d/ds at s = 1.1 is the engine's value / 1.1
    th2 = theta0.copy()
    th2[2 * NZ] = 1.1
    _, g2 = prob.evaluate(th2, gradient=True)
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    Q.porosity = P.porosity * 1.1
    E = Engine(Q)
    try:
        E.load_particles(*release(P))
        out = E.adjoint(WINDOWS, obs, prob.target, (1.5, 1.5, 1.0))
    finally:
        E.close()
    assert g2[2 * NZ] == pytest.approx(out["grad_porosity_scale"] / 1.1, rel=1e-12)


def _gradient(P, zones, aL, aT, dm, target, opts=None, state=None, windows=WINDOWS):
    """Gradient of the misfit with the sweep options `opts`; `state` (a
    particles() dict at time t0) starts the run from that state instead."""
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    E = Engine(Q)
    try:
        if opts:
            E.adjoint_options(**opts)
        if state is None:
            E.load_particles(*release(P))
        else:
            t0, g = state
            E.load_particles(g["cell"], g["lx"], g["ly"], g["lz"], t=g["t"],
                             status=g["status"], rng=g["rng"])
            E.set_time(t0)
        obs = observations(P)
        out = E.adjoint(windows, obs, target, (1.5, 1.5, 1.0), gradient=True)
        if opts and opts.get("diagnostics"):
            out["stats"] = E.adjoint_stats()
    finally:
        E.close()
    return out


def _same_gradient(a, b, rtol=1e-10):
    """Per-cell gradients are sums of FP64 atomics: equal to rounding."""
    for k in ("grad_al", "grad_at", "grad_dm"):
        scale = np.abs(a[k]).max() + 1e-300
        np.testing.assert_allclose(a[k], b[k], rtol=0, atol=rtol * scale, err_msg=k)
    assert a["grad_porosity_scale"] == pytest.approx(b["grad_porosity_scale"], rel=rtol, abs=1e-300)


def test_sweep_diagnostics_and_graze_cut():
    """The diagnostics change no gradient and classify every exit by its
    kappa; graze_cut drops exactly the exits below it, and a cut below
    every exit's kappa leaves the gradient as it is; frozen_noise keeps the
    misfit and changes the gradient of a walk with exits."""
    P, zones, aL, aT, dm = base()
    target = 0.7 * run(P, zones, aL, aT, dm)["conc"][::-1]
    plain = _gradient(P, zones, aL, aT, dm, target)
    d = _gradient(P, zones, aL, aT, dm, target, {"diagnostics": True})
    _same_gradient(plain, d)
    S = d["stats"]
    print("sweep stats:", S, flush=True)
    kinds = S["kinds"]
    exits = kinds["exit"]["iterations"] + kinds["grazing_exit"]["iterations"]
    bins = list(S["exit_kappa"].values())
    assert exits > 0 and sum(bins) == exits
    assert kinds["substep"]["iterations"] > 0
    assert S["exits_cut"] == 0
    assert 0 < S["particles"] <= 400
    for k in kinds.values():
        assert k["log10_growth_max"] >= 0 and k["growth_over_10"] <= k["iterations"]
    # a cut at 1e-2 drops the exits of the five lowest kappa bins
    c = _gradient(P, zones, aL, aT, dm, target, {"graze_cut": 1e-2, "diagnostics": True})
    assert c["stats"]["exits_cut"] == sum(bins[:5])
    if sum(bins[:5]) == 0:
        _same_gradient(plain, c)
    # a cut at 1 drops every exit whose root terms cancel at all (kappa < 1);
    # kappa is exactly 1 where 2 A m and B have the same sign
    full = _gradient(P, zones, aL, aT, dm, target, {"graze_cut": 1.0, "diagnostics": True})
    assert sum(bins[:6]) <= full["stats"]["exits_cut"] <= exits
    assert np.isfinite(full["grad_al"]).all()
    if full["stats"]["exits_cut"] > 0:
        assert not np.allclose(full["grad_al"], plain["grad_al"], rtol=1e-9, atol=0)
    # frozen noise: the sub-steps' random displacements lose their duration
    # derivative, so a walk with cell exits gets another (finite) gradient
    fz = _gradient(P, zones, aL, aT, dm, target, {"frozen_noise": True})
    assert fz["misfit"] == pytest.approx(plain["misfit"], rel=1e-12)
    assert np.isfinite(fz["grad_al"]).all() and np.isfinite(fz["grad_porosity_scale"])
    assert not np.allclose(fz["grad_al"], plain["grad_al"], rtol=1e-6, atol=0)
    # the lowest non-empty bin's upper edge leaves those exits uncut
    edges = (1e-10, 1e-8, 1e-6, 1e-4, 1e-2, 1e-1)
    low = next(i for i, b in enumerate(bins) if b > 0)
    if low > 0:
        below = _gradient(P, zones, aL, aT, dm, target, {"graze_cut": edges[low - 1]})
        _same_gradient(plain, below)


def test_sweep_truncated_to_last_window():
    """reverse_windows = 1 is the gradient of the last window alone: that
    of a run started from the state at the last window's start (the
    checkpoint, random streams included)."""
    P, zones, aL, aT, dm = base()
    target = 0.7 * run(P, zones, aL, aT, dm)["conc"][::-1]
    last = _gradient(P, zones, aL, aT, dm, target, {"reverse_windows": 1})
    every = _gradient(P, zones, aL, aT, dm, target, {"reverse_windows": len(WINDOWS)})
    _same_gradient(_gradient(P, zones, aL, aT, dm, target), every)
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    E = Engine(Q)
    try:
        E.load_particles(*release(P))
        for t in WINDOWS[:-1]:
            E.advance(t)
        state = (WINDOWS[-2], E.particles())
    finally:
        E.close()
    alone = _gradient(P, zones, aL, aT, dm, target, state=state, windows=WINDOWS[-1:])
    assert alone["misfit"] == pytest.approx(last["misfit"], rel=1e-12)
    _same_gradient(alone, last)
    assert not np.allclose(last["grad_al"], every["grad_al"])


def test_tangent_restart_is_the_window_local_sensitivity():
    """tangent_windows = 1: window w contributes lambda_T . dX(t_w+1)/dtheta
    with the parameters perturbed during window w only, lambda_T the
    final-time derivative of the misfit with respect to each particle's
    global position.  Window 2 of 3 (the difference of the sweeps of the
    last two and the last window) against central differences of the
    particles' positions at its end, common random numbers, the
    parameters changed at its start; the final-time gradient from the
    misfit formula evaluated here."""
    P, zones, aL, aT, dm = base()
    obs = observations(P)
    bw = np.array([1.5, 1.5, 1.0])
    target = 0.7 * run(P, zones, aL, aT, dm)["conc"][::-1]
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    E = Engine(Q)
    try:
        E.load_particles(*release(P))
        for t in WINDOWS:
            E.advance(t)
        X = np.stack(E.global_xyz(), axis=1)
        on = E.particles()["status"] == 1
    finally:
        E.close()
    u = (X[on, None, :] - obs[None, :, :]) / bw            # (p, o, d)
    K = np.prod(np.exp(-0.5 * u ** 2) / (bw * np.sqrt(2 * np.pi)), axis=2)
    resid = K.sum(axis=0) - target
    gX = np.zeros_like(X)
    gX[on] = np.einsum("o,po,pod->pd", 2 * resid, K, -u / bw)

    adj = {}
    for rw in (1, 2):
        adj[rw] = _gradient(P, zones, aL, aT, dm, target,
                            {"tangent_windows": 1, "reverse_windows": rw})
    assert adj[1]["misfit"] == pytest.approx((resid ** 2).sum(), rel=1e-10)
    last = _gradient(P, zones, aL, aT, dm, target, {"reverse_windows": 1})
    _same_gradient(adj[1], last)

    def window2_positions(aL2, aT2, dm2):
        E = Engine(Q)
        try:
            E.load_particles(*release(P))
            E.advance(WINDOWS[0])
            E.set_dispersion(problems.zoned_dispersion(P, zones, aL2, aT2, dm2))
            E.advance(WINDOWS[1])
            g = E.particles()
            return np.stack(E.global_xyz(), axis=1), (g["cell"], g["status"], g["substeps"])
        finally:
            E.close()

    checked = 0
    params = [("aL", z) for z in range(NZ)] + [("aT", 0), ("dm", 0)]
    for name, z in params:
        key = {"aL": "grad_al", "aT": "grad_at", "dm": "grad_dm"}[name]
        cells = zones == z if name != "dm" else np.ones(len(zones), bool)
        g = adj[2][key][cells].sum() - adj[1][key][cells].sum()
        for rel in (1e-6, 1e-7, 1e-8):
            runs = []
            for sg in (1.0, -1.0):
                a2, t2, d2 = aL.copy(), aT.copy(), dm
                if name == "aL":
                    h = rel * aL[z]
                    a2[z] += sg * h
                elif name == "aT":
                    h = rel * aT[z]
                    t2[z] += sg * h
                else:
                    h = rel * dm
                    d2 = dm + sg * h
                runs.append(window2_positions(a2, t2, d2))
            if not all(np.array_equal(a, b) for a, b in zip(runs[0][1], runs[1][1])):
                continue
            fd = float((gX * (runs[0][0] - runs[1][0])).sum() / (2 * h))
            print(f"  window 2 {name}{z} rel {rel:g}: adjoint {g:.9e} fd {fd:.9e}", flush=True)
            if abs(g - fd) <= 1e-4 * abs(fd) + 1e-12 * np.abs(gX).sum() / h:
                checked += 1
                break
    assert checked >= len(params) - 1
    every = _gradient(P, zones, aL, aT, dm, target, {"tangent_windows": len(WINDOWS)})
    _same_gradient(_gradient(P, zones, aL, aT, dm, target), every)


def _dual_misfit(Q, rel, obs, tgt, bw, tangents, windows=WINDOWS):
    """The misfit and its derivatives along the tangent directions from the
    dual-number oracle: final global positions with tangents, then the
    Gaussian-smoothed concentrations of rwpt_adjoint differentiated by hand
    (d N(z)/dz = -z N(z))."""
    from oracle import oracle_dual as D
    DP = D.DualProblem(Q, tangents)
    st = D.new_state(*rel)
    t0 = 0.0
    for t1 in windows:
        assert D.advance(DP, st, t0, t1) == 0
        t0 = t1
    act = st["status"] == 1
    g = D.global_xyz(DP, st["cell"][act], st["lx"][act], st["ly"][act], st["lz"][act])
    X = np.stack([a[:, 0] for a in g], axis=1)                  # (n, 3)
    dX = np.stack([a[:, 1:] for a in g], axis=1)                # (n, 3, ND)
    h = np.asarray(bw, float)
    J, dJ = 0.0, np.zeros(D.ND)
    for o in range(len(obs)):
        z = (X - obs[o]) / h
        K = np.prod(np.exp(-0.5 * z * z) / (np.sqrt(2.0 * np.pi) * h), axis=1)
        dK = K[:, None] * np.einsum("nd,ndk->nk", -z / h, dX)
        c, dc = K.sum(), dK.sum(axis=0)
        J += (c - tgt[o]) ** 2
        dJ += 2.0 * (c - tgt[o]) * dc
    return J, dJ


def test_gradient_equals_forward_mode_duals():
    """The FP64 reference gradient: forward-mode dual numbers through the
    oracle's cell records and walk (oracle/rwpt_oracle_dual.cpp, pinned
    against central differences of the oracle by test_oracle_dual.py) give
    the misfit's derivatives along random per-cell directions of aL, aT and
    Dm and along the porosity scale.  The device's reverse sweep must give
    the same numbers as the dot products of its per-cell gradients with the
    directions (its drift-term tangents are FP32: 1e-6 relative)."""
    P, zones, aL, aT, dm = base()
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    obs = observations(P)
    bw = (1.5, 1.5, 1.0)
    first = run(P, zones, aL, aT, dm)
    tgt = 0.5 * first["conc"]
    out = run(P, zones, aL, aT, dm, gradient=True, target=tgt)
    g = np.random.default_rng(3)
    v = [g.uniform(-1.0, 1.0, P.ncell) for _ in range(3)]
    tang = {k: np.zeros((P.ncell, 4)) for k in ("aL", "aT", "dmeff", "porosity")}
    tang["aL"][:, 0], tang["aT"][:, 1], tang["dmeff"][:, 2] = v
    tang["porosity"][:, 3] = Q.porosity
    J, dJ = _dual_misfit(Q, release(P), obs, tgt, bw, tang)
    assert J == pytest.approx(out["misfit"], rel=1e-12)
    want = np.array([np.dot(out["grad_al"], v[0]), np.dot(out["grad_at"], v[1]),
                     np.dot(out["grad_dm"], v[2]), out["grad_porosity_scale"]])
    print("adjoint vs duals:", want, dJ, flush=True)
    np.testing.assert_allclose(want, dJ, rtol=1e-6, atol=0)
    assert np.all(np.abs(dJ) > 0)


def test_long_horizon_gradient_equals_duals():
    """Over 30 days (60 RW steps, 6 checkpoint windows, many cell exits per
    particle) the reverse sweep across window boundaries still gives the
    forward-mode derivatives."""
    P, zones, aL, aT, dm = base()
    Q = problems.zoned_dispersion(P, zones, aL, aT, dm)
    windows = (5.0, 10.0, 15.0, 20.0, 25.0, 30.0)
    obs = problems.observation_points(P, 6, 3, (8, 22), (7, 14))
    bw = (1.5, 1.5, 1.0)
    E = Engine(Q)
    try:
        E.load_particles(*release(P))
        tgt = 0.5 * E.adjoint(windows, obs, np.zeros(len(obs)), bw, gradient=False)["conc"]
    finally:
        E.close()
    E = Engine(Q)
    try:
        E.load_particles(*release(P))
        out = E.adjoint(windows, obs, tgt, bw, gradient=True)
    finally:
        E.close()
    g = np.random.default_rng(4)
    v = [g.uniform(-1.0, 1.0, P.ncell) for _ in range(3)]
    tang = {k: np.zeros((P.ncell, 4)) for k in ("aL", "aT", "dmeff", "porosity")}
    tang["aL"][:, 0], tang["aT"][:, 1], tang["dmeff"][:, 2] = v
    tang["porosity"][:, 3] = Q.porosity
    J, dJ = _dual_misfit(Q, release(P), obs, tgt, bw, tang, windows=windows)
    assert J == pytest.approx(out["misfit"], rel=1e-12)
    want = np.array([np.dot(out["grad_al"], v[0]), np.dot(out["grad_at"], v[1]),
                     np.dot(out["grad_dm"], v[2]), out["grad_porosity_scale"]])
    print("long horizon, adjoint vs duals:", want, dJ, flush=True)
    np.testing.assert_allclose(want, dJ, rtol=1e-5, atol=0)
