"""GPU parity of the gpkde reconstruction (rwpt_kde) against the oracle.

Unit-mass histogram counts are integers and must agree exactly; weighted
counts sum each bin in the reference's particle order and agree to
rounding of the effective-mass round trip the reference performs (1e-14).
The density is the reference's scatter restated as three separable passes,
so it agrees to summation order (1e-10 relative, 1e-13 of the maximum).
"""
import numpy as np
import pytest

from oracle import oracle as O
from rwpt import Engine, problems

pytestmark = pytest.mark.gpu


def _walked(P, where, n, t_end, steps):
    if where is None:
        rel = problems.release_in_block(P, n, (3, 8), (6, 15), (1, P.nlay), seed=3)
    else:
        rel = problems.release_in_cell(P, n, *where, seed=3, lz=0.5)
    E = Engine(P)
    E.load_particles(*rel)
    for k in range(1, steps + 1):
        E.advance(t_end * k / steps)
    return E


def _grid(P):
    nb = (P.ncol, P.nrow, P.nlay)
    return (0.0, 0.0, float(P.bot.min())), (float(P.delr[0]), float(P.delc[0]),
                                            float(P.top[0] - P.bot[0])), nb


@pytest.mark.parametrize("case", ["cfg0", "hetero3d", "hetero3d_weighted"])
def test_kde_matches_oracle(case):
    if case == "cfg0":
        P = problems.config0(ts_params=(0.05, 0.0))
        E = _walked(P, (1, 25, 10), 4000, 10.0, 4)
    else:
        P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
        P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
        P.ts_params = (0.5, 0.0)
        E = _walked(P, None, 4000, 4.0, 2)
    try:
        origin, binsize, nbins = _grid(P)
        g = E.particles()
        mass = None
        if case.endswith("weighted"):
            mass = np.random.default_rng(5).uniform(0.2, 3.0, g["cell"].size)
        dens, hist = E.kde(origin, binsize, nbins, mode=1, mass=mass)
        info = E.kde_info()
        raw, hist0 = E.kde(origin, binsize, nbins, mode=0, mass=mass)
    finally:
        E.close()
    act = g["status"] == 1
    gx, gy, gz = O.global_xyz(P, g["cell"], g["lx"], g["ly"], g["lz"])
    od, oc, oh = O.kde_density(gx[act], gy[act], gz[act], origin, binsize, nbins,
                               None if mass is None else mass[act])
    if mass is None:
        assert np.array_equal(hist, oc)
    else:
        np.testing.assert_allclose(hist, oc, rtol=1e-14, atol=0)
    assert abs(info["bandwidth"] - oh[0]) <= 1e-12 * oh[0]
    np.testing.assert_allclose(dens, od, rtol=1e-10, atol=1e-13 * od.max())
    np.testing.assert_array_equal(hist0, hist)
    vol = np.prod(binsize)
    np.testing.assert_allclose(raw, hist / vol, rtol=1e-15)
    assert abs(dens.sum() * vol - hist.sum()) <= 1e-9 * hist.sum()


@pytest.mark.parametrize("n,spread", [(1, 0.0), (3, 4.5), (8, 4.0)])
def test_kernel_wider_than_the_grid(n, spread):
    """Kernels whose half-width exceeds a grid dimension (3 layers, h up to
    2 bins): the reference folds the part below the grid once and then the
    part above it once, and drops what lands outside after that (its density
    loses mass); a single particle takes the zero-variance bandwidth."""
    P = problems.heterogeneous_flow(14, 10, 3, corr_len=3.0, seed=11)
    g = np.random.default_rng(n)
    s = int(spread)
    cell = P.cell_number(g.integers(1, 4, n), g.integers(5 - s, 6 + s, n),
                         g.integers(7 - s, 8 + s, n))
    if n == 1:
        cell = P.cell_number(np.array([3]), np.array([10]), np.array([10]))
    lx, ly, lz = g.uniform(0.0, 1.0, n), g.uniform(0.0, 1.0, n), g.uniform(0.0, 0.05, n)
    origin, binsize, nbins = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (14, 10, 3)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        dens, hist = E.kde(origin, binsize, nbins, mode=1, mass=np.ones(n))
        h = E.kde_info()["bandwidth"]
    finally:
        E.close()
    gx, gy, gz = O.global_xyz(P, cell.astype(np.int32), lx, ly, lz)
    od, oc, oh = O.kde_density(gx, gy, gz, origin, binsize, nbins, np.ones(n))
    assert np.array_equal(hist, oc)
    assert abs(h - oh[0]) <= 1e-12 * oh[0] and 3 * oh[0] > 3
    np.testing.assert_allclose(dens, od, rtol=0, atol=1e-12 * od.max())
    # the reference's folds drop mass when the kernel is wider than the grid
    assert dens.sum() < 0.9999 * n


def test_global_z_uses_the_saturated_top():
    """A convertible top layer whose saturated top lies below the cell top
    (rwpt_set_saturated_top): the global z and the histogram follow
    ConvertToModelXYZ with Grid%SaturatedTop."""
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    sat = P.top.copy()
    top_layer = np.arange(P.ncol * P.nrow)
    sat[top_layer] = P.bot[top_layer] + (P.top - P.bot)[top_layer] * \
        np.random.default_rng(2).uniform(0.2, 0.9, top_layer.size)
    P.saturated_top = sat
    E = _walked(P, None, 4000, 4.0, 2)
    try:
        g = E.particles()
        gx, gy, gz = E.global_xyz()
        origin, binsize, nbins = _grid(P)
        _, hist = E.kde(origin, binsize, nbins, mode=0)
        E.set_saturated_top(None)
        gz_top = E.global_xyz()[2]
    finally:
        E.close()
    ox, oy, oz = O.global_xyz(P, g["cell"], g["lx"], g["ly"], g["lz"])
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy) and np.array_equal(gz, oz)
    in_top = g["cell"] <= P.ncol * P.nrow
    assert in_top.sum() > 100
    assert (gz[in_top] < gz_top[in_top]).mean() > 0.9
    assert np.array_equal(gz[~in_top], gz_top[~in_top])
    act = g["status"] == 1
    assert np.array_equal(hist, O.histogram(ox[act], oy[act], oz[act], origin, binsize,
                                            nbins))


def test_kde_full_size_mass_conservation():
    P = problems.config1()
    n = 1_000_000
    rel = problems.release_in_block(P, n, (1, 10), (76, 125), (1, 20))
    E = Engine(P)
    try:
        E.load_particles(*rel)
        E.advance(25.0)
        dens, hist = E.kde((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (200, 200, 20), mode=1)
        info = E.kde_info()
        st = E.particles()["status"]
        dens2, hist2 = E.kde((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (200, 200, 20), mode=1)
    finally:
        E.close()
    # active particles on the outer face (local coordinate 1 of the last
    # column/row/layer) bin outside the grid, as in the reference's histogram
    act = (st == 1).sum()
    assert act == info["points"] and act - 100 <= hist.sum() <= act
    assert np.all(hist == np.round(hist))
    assert abs(dens.sum() - hist.sum()) <= 1e-9 * hist.sum()
    assert np.array_equal(dens, dens2) and np.array_equal(hist, hist2)
    assert info["bandwidth"] > 1.0


@pytest.mark.parametrize("unit_volume", [False, True])
def test_kde_bins_unlike_cells(unit_volume):
    """Bins of 0.5 x 2 x 1.5 m over the 1 m flow cells, a scaling factor,
    and the unit-volume form MPathRW requests."""
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    E = _walked(P, None, 5000, 4.0, 2)
    origin, binsize, nbins = (0.0, 0.0, 0.0), (0.5, 2.0, 1.5), (48, 10, 4)
    sf = 0.37
    try:
        g = E.particles()
        dens, hist = E.kde(origin, binsize, nbins, mode=1, scaling_factor=sf,
                           unit_volume=unit_volume)
    finally:
        E.close()
    act = g["status"] == 1
    gx, gy, gz = O.global_xyz(P, g["cell"], g["lx"], g["ly"], g["lz"])
    od, oc, _ = O.kde_density(gx[act], gy[act], gz[act], origin, binsize, nbins)
    assert np.array_equal(hist, oc)
    want = od * sf * (np.prod(binsize) if unit_volume else 1.0)
    np.testing.assert_allclose(dens, want, rtol=1e-10, atol=1e-13 * want.max())
