"""GPU parity of the gpkde bandwidth optimisation (nOptimizationLoops > 0).

The device loop is compared with the numpy restatement of
GridProjectedKDE%ComputeDensityOptimization (itself pinned loop by loop to
the compiled reference) on the same histogram, and with the compiled
reference directly where its binary is built.  The scatters add with
atomics, so the densities agree to summation order: 1e-9 of the maximum
after ten loops of pow/sqrt/erf arithmetic on the gathered sums.
"""
import numpy as np
import pytest

from oracle import gpkde_opt, refrun
from oracle import oracle as O
from rwpt import Engine, RwptError, problems

pytestmark = pytest.mark.gpu


def unit_grid(nc, nr, nl):
    n = nc * nr * nl
    k = np.repeat(np.arange(nl), nr * nc)
    top = (nl - k).astype(float)
    return problems.Problem(ncol=nc, nrow=nr, nlay=nl, delr=np.ones(nc), delc=np.ones(nr),
                            top=top, bot=top - 1.0, frf=np.zeros(n), fff=np.zeros(n),
                            flf=np.zeros(n), porosity=np.full(n, 0.3),
                            dim_mask=(1, 1, 1 if nl > 1 else 0), ts_params=(0.1, 0.0))


def cloud(P, n, seed):
    """Points of a skewed Gaussian cloud, as (cell, local x, y, z)."""
    g = np.random.default_rng(seed)
    x = np.clip(g.normal(P.ncol * 0.5, P.ncol * 0.12, n), 0.01, P.ncol - 0.01)
    y = np.clip(g.gamma(4.0, P.nrow * 0.06, n) + 0.1 * P.nrow, 0.01, P.nrow - 0.01)
    if P.nlay > 1:
        z = np.clip(g.normal(P.nlay * 0.5, P.nlay * 0.15, n), 0.01, P.nlay - 0.01)
    else:
        z = g.uniform(0.01, 0.99, n)
    j = np.floor(x).astype(int)
    i = np.floor(P.nrow - y).astype(int)
    k = np.floor(P.nlay - z).astype(int)
    cell = (k * P.nrow * P.ncol + i * P.ncol + j + 1).astype(np.int32)
    return cell, x - j, y - (P.nrow - i - 1), z - (P.nlay - k - 1)


CASES = {
    # name: grid, bins (size, count), loops, weighted, isotropic, unit volume, sf
    "2d": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 10, False, False, False, 1.0),
    "2d_weighted": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, True, False, False, 1.0),
    "2d_iso": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, False, True, False, 1.0),
    "2d_bins": ((60, 40, 1), ((2.0, 0.5, 1.0), (30, 80, 1)), 10, False, False, True, 0.3),
    "3d": ((24, 20, 12), ((1.0, 1.0, 1.0), (24, 20, 12)), 4, False, False, False, 1.0),
    "3d_weighted_iso": ((24, 20, 12), ((1.0, 1.0, 1.0), (24, 20, 12)), 3, True, True,
                        True, 2.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_adaptive_density(name):
    grid, (bs, nb), loops, weighted, iso, uv, sf = CASES[name]
    P = unit_grid(*grid)
    cell, lx, ly, lz = cloud(P, 20000, 11)
    mass = np.random.default_rng(2).lognormal(0.0, 0.4, cell.size) if weighted else None
    origin = (0.0, 0.0, 0.0)
    print(f"gpkde opt {name}: grid {grid} bins {nb} loops {loops}", flush=True)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        dens, hist = E.kde(origin, bs, nb, mode=1, mass=mass, scaling_factor=sf,
                           unit_volume=uv, n_opt_loops=loops, isotropic=iso)
        info = E.kde_info()
    finally:
        E.close()
    gx, gy, gz = O.global_xyz(P, cell, lx, ly, lz)
    od, oc, oh = O.kde_density(gx, gy, gz, origin, bs, nb, mass)
    assert abs(info["bandwidth"] - oh[0]) <= 1e-12 * oh[0]
    if mass is None:
        assert np.array_equal(hist, oc)
        counts, eff, npts = oc, 1.0, oc.sum()
    else:
        eff = mass.sum() / (mass.sum() ** 2 / (mass * mass).sum())
        counts, npts = oc / eff, float(mass.size)
    want, oloops = gpkde_opt.optimize(counts, nb, bs, oh, npts, loops, isotropic=iso)
    want = want * eff * sf * (np.prod(bs) if uv else 1.0)
    assert info["loops"] == oloops
    np.testing.assert_allclose(dens, want, rtol=0, atol=1e-9 * want.max())
    if refrun.gpkde_available():
        dr, _, _ = refrun.gpkde(gx, gy, gz, origin, bs, nb, weights=mass, n_opt_loops=loops,
                                isotropic=iso, unit_volume=uv, scaling_factor=sf)
        np.testing.assert_allclose(dens, dr, rtol=0, atol=1e-9 * dr.max())


def test_one_dimensional_loop_unsupported():
    P = unit_grid(60, 40, 1)
    cell, lx, ly, lz = cloud(P, 2000, 4)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        with pytest.raises(RwptError):
            E.kde((0.0, 0.0, 0.0), (1.0, 40.0, 1.0), (60, 1, 1), mode=1, n_opt_loops=2)
    finally:
        E.close()
