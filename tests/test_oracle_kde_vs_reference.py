"""The oracle's gpkde restatement against the reference's own
GridProjectedKDE (compiled from its sources by oracle/ref/Makefile, driven
as MPathRW.F90:556-574 and :1626-1642 drive it) with nOptimizationLoops = 0
and the brute-force kernel.

Histogram counts and the Silverman bandwidth agree exactly; the density is
the same scatter of erf-integrated kernels summed in another order, so it
agrees to 1e-13 of its maximum.  With unitVolume the reference multiplies
the density back by the bin volume (GridProjectedKDE.F90:4292-4300).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import refrun

pytestmark = pytest.mark.skipif(not refrun.available(),
                                reason="reference driver not built (oracle/ref)")


def _cloud(case, n=3000, seed=0):
    g = np.random.default_rng(seed)
    if case.startswith("2d"):
        nb, bs = (40, 30, 1), (1.0, 1.0, 1.0)
        pts = (g.normal(20, 4, n), g.normal(15, 3, n), g.uniform(0, 1, n))
    else:
        nb, bs = (20, 16, 12), (0.5, 2.0, 1.5)
        pts = (g.normal(5, 2, n), g.normal(16, 6, n), g.normal(9, 3, n))
    w = g.uniform(0.2, 3.0, n) if "w" in case.split("_")[1:] else None
    return pts, nb, bs, w


@pytest.mark.parametrize("case", ["2d", "3d", "2d_w", "3d_w"])
@pytest.mark.parametrize("unit_volume", [False, True])
def test_kde_oracle_matches_reference(case, unit_volume):
    (x, y, z), nb, bs, w = _cloud(case)
    sf = 0.37
    d_ref, c_ref, h_ref = refrun.gpkde(x, y, z, (0.0, 0.0, 0.0), bs, nb, weights=w,
                                       unit_volume=unit_volume, scaling_factor=sf)
    d_or, c_or, h_or = O.kde_density(x, y, z, (0.0, 0.0, 0.0), bs, nb, w)
    np.testing.assert_array_equal(h_ref, h_or)
    if w is None:
        np.testing.assert_array_equal(c_ref, c_or * sf)
    else:
        np.testing.assert_allclose(c_ref, c_or * sf, rtol=1e-14, atol=0)
    want = d_or * sf * (np.prod(bs) if unit_volume else 1.0)
    np.testing.assert_allclose(d_ref, want, rtol=0, atol=1e-13 * want.max())
    assert want.max() > 0


def test_kde_single_point_cloud_matches_reference():
    """A zero standard deviation falls back to a bandwidth of twice the
    bin size (GridProjectedKDE.F90:4043-4054)."""
    x, y, z = np.full(50, 10.3), np.full(50, 7.6), np.full(50, 0.5)
    d_ref, c_ref, h_ref = refrun.gpkde(x, y, z, (0, 0, 0), (1, 1, 1), (20, 15, 1))
    d_or, c_or, h_or = O.kde_density(x, y, z, (0, 0, 0), (1, 1, 1), (20, 15, 1))
    np.testing.assert_array_equal(h_ref, h_or)
    np.testing.assert_array_equal(c_ref, c_or)
    np.testing.assert_allclose(d_ref, d_or, rtol=0, atol=1e-13 * d_or.max())


def test_kde_single_point_cloud_anisotropic_bins():
    """The fallback uses the bin distance, the bin volume to the power one
    over the number of dimensions with a bin size (Histogram.F90:154-155),
    three here although the reconstruction is two-dimensional.  The
    coordinates are exact binary fractions, so the deviations are zero."""
    x, y, z = np.full(64, 10.25), np.full(64, 7.625), np.full(64, 0.5)
    bs, nb = (2.0, 0.5, 1.5), (20, 40, 1)
    d_ref, c_ref, h_ref = refrun.gpkde(x, y, z, (0, 0, 0), bs, nb, unit_volume=False)
    d_or, c_or, h_or = O.kde_density(x, y, z, (0, 0, 0), bs, nb)
    np.testing.assert_allclose(h_or[:2], 2.0 * 1.5 ** (1.0 / 3.0), rtol=1e-15)
    np.testing.assert_array_equal(h_ref, h_or)
    np.testing.assert_array_equal(c_ref, c_or)
    np.testing.assert_allclose(d_ref, d_or, rtol=0, atol=1e-13 * d_or.max())


@pytest.mark.parametrize("zbin", [1.0, 0.0])
def test_kde_single_bin_dimension(zbin):
    """A dimension of one bin with a bin size drops the points outside it
    (Histogram.F90:121, 306-317); a zero bin size compresses it."""
    g = np.random.default_rng(4)
    n = 4000
    x, y, z = g.normal(20, 4, n), g.normal(15, 3, n), g.uniform(-0.6, 1.6, n)
    nb, bs = (40, 30, 1), (1.0, 1.0, zbin)
    d_ref, c_ref, h_ref = refrun.gpkde(x, y, z, (0, 0, 0), bs, nb, unit_volume=False)
    d_or, c_or, h_or = O.kde_density(x, y, z, (0, 0, 0), bs, nb)
    inside = (z >= 0) & (z < 1) if zbin else np.ones(n, bool)
    assert c_or.sum() <= inside.sum()
    np.testing.assert_array_equal(c_ref, c_or)
    np.testing.assert_array_equal(h_ref, h_or)
    np.testing.assert_allclose(d_ref, d_or, rtol=0, atol=1e-13 * d_or.max())


@pytest.mark.parametrize("sel,factor", [(0, 1.3), (1, 1.7), (1, 0.6)])
@pytest.mark.parametrize("case", ["2d", "3d_w"])
def test_kde_initial_smoothing_selection(sel, factor, case):
    """initialSmoothingSelection 0 scales Silverman's bandwidth by a factor
    other than one; 1 takes the factor times the bin distance
    (GridProjectedKDE.F90:631-650, 4109-4117)."""
    (x, y, z), nb, bs, w = _cloud(case)
    d_ref, c_ref, h_ref = refrun.gpkde(x, y, z, (0.0, 0.0, 0.0), bs, nb, weights=w,
                                       unit_volume=False, smoothing_selection=sel,
                                       smoothing_factor=factor)
    d_or, c_or, h_or = O.kde_density(x, y, z, (0.0, 0.0, 0.0), bs, nb, w, sel, factor)
    np.testing.assert_allclose(h_ref, h_or, rtol=1e-15, atol=0)
    np.testing.assert_allclose(d_ref, d_or, rtol=0, atol=1e-13 * d_or.max())


@pytest.mark.parametrize("n,spread", [(1, 0.0), (3, 4.5), (6, 4.5)])
def test_kde_kernel_wider_than_the_grid(n, spread):
    """Three layers of bins and kernels up to 6 bins wide (a single point
    takes the zero-variance bandwidth): the reference folds a kernel's
    entries below the grid once and then those above it once
    (KernelMultiGaussian.f90:303-390) and drops what still lies outside, so
    the density loses mass; the oracle's fold does the same."""
    g = np.random.default_rng(n)
    x, y = 7 + g.uniform(-spread, spread, n), 5 + g.uniform(-spread, spread, n)
    z = g.uniform(0.0, 3.0, n) if n > 1 else np.array([0.0083])
    nb, bs = (14, 10, 3), (1.0, 1.0, 1.0)
    d_ref, c_ref, h_ref = refrun.gpkde(x, y, z, (0.0, 0.0, 0.0), bs, nb, weights=np.ones(n),
                                       unit_volume=True)
    d_or, c_or, h_or = O.kde_density(x, y, z, (0.0, 0.0, 0.0), bs, nb, np.ones(n))
    assert np.array_equal(c_or, c_ref)
    np.testing.assert_array_equal(h_or, h_ref)
    assert 3 * h_ref[2] > 3.0
    np.testing.assert_allclose(d_or, d_ref, rtol=0, atol=1e-13 * d_ref.max())
    assert d_ref.sum() < 0.9999 * n
