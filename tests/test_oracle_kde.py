"""Pins of the oracle's gpkde density (GridProjectedKDE.F90 ComputeDensity
with nOptimizationLoops = 0 and the brute-force kernel).

* the reference scatters each active bin's erf-integrated multi-Gaussian and
  folds what falls outside the grid back across the boundary
  (KernelMultiGaussian.f90:303-390); for a uniform bandwidth that is a
  separable convolution with half-sample symmetric padding, computed here
  independently with scipy.ndimage.convolve1d(mode="reflect");
* the bandwidth is Silverman's rule sigma (4/((d+2)n))^(1/(d+4)) with the
  geometric-mean sigma of the active dimensions;
* the reflected estimate conserves mass: sum(density) * binVolume equals
  the histogram total;
* histogram counts are integers for unit masses and match numpy's
  histogramdd with the reference's floor((x-origin)/bin)+1 binning.
"""
import numpy as np
import pytest
from scipy import ndimage
from scipy.special import erf

from oracle import oracle as O


def _kernel1d(h, b):
    m = int(np.ceil(3 * h / b))
    i = np.arange(-m, m + 1)
    e = erf((np.abs(i) + 0.5) / (h / b * np.sqrt(2))) - erf((np.abs(i) - 0.5) / (h / b * np.sqrt(2)))
    return e / e.sum()


def _reference_like(x, y, z, origin, binsize, nbins, w=None):
    pts = [x, y, z]
    act = [nb > 1 for nb in nbins]
    nd = sum(act)
    edges = [origin[d] + binsize[d] * np.arange(nbins[d] + 1) for d in range(3)]
    sample = np.stack(pts, 1)
    ww = np.ones(x.size) if w is None else w
    idx = [np.floor((pts[d] - origin[d]) / binsize[d]).astype(int) if act[d]
           else np.zeros(x.size, int) for d in range(3)]
    inside = np.ones(x.size, bool)
    for d in range(3):
        if act[d]:
            inside &= (idx[d] >= 0) & (idx[d] < nbins[d])
    H = np.zeros(nbins[::-1])  # (z, y, x)
    np.add.at(H, (idx[2][inside], idx[1][inside], idx[0][inside]), ww[inside])
    n_points = x.size if w is not None else int(inside.sum())
    sd = [np.sqrt(((pts[d] - pts[d].mean()) ** 2).mean()) for d in range(3)]
    s = np.prod([sd[d] for d in range(3) if act[d]]) ** (1.0 / nd)
    h = s * (4.0 / ((nd + 2.0) * n_points)) ** (1.0 / (nd + 4.0))
    D = H.copy()
    for d in range(3):
        if act[d]:
            D = ndimage.convolve1d(D, _kernel1d(h, binsize[d]), axis=2 - d, mode="reflect")
    vol = np.prod(binsize)
    return D / vol, H, h


@pytest.mark.parametrize("case", ["2d", "3d", "3d_weighted_edges"])
def test_density_matches_separable_reflect_convolution(case):
    rng = np.random.default_rng(4)
    if case == "2d":
        n, origin, binsize, nbins = 3000, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (40, 25, 1)
        x = rng.normal(12.0, 3.0, n); y = rng.normal(6.0, 2.0, n); z = rng.uniform(0, 1, n)
        w = None
    elif case == "3d":
        n, origin, binsize, nbins = 4000, (0.0, 0.0, 0.0), (1.0, 0.5, 2.0), (30, 30, 8)
        x = rng.normal(15, 4, n); y = rng.normal(7, 2, n); z = rng.normal(8, 3, n)
        w = None
    else:  # heavy reflection near every face, some points outside
        n, origin, binsize, nbins = 3000, (-1.0, 2.0, 0.0), (0.5, 1.0, 1.0), (24, 18, 10)
        x = rng.uniform(-2, 12, n); y = rng.uniform(1, 21, n); z = rng.uniform(-1, 11, n)
        w = rng.uniform(0.5, 2.0, n)
    dens, counts, h = O.kde_density(x, y, z, origin, binsize, nbins, w)
    D, H, hh = _reference_like(x, y, z, origin, binsize, nbins, w)
    assert abs(h[0] - hh) <= 1e-13 * hh
    np.testing.assert_allclose(counts.reshape(nbins[::-1]), H, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dens.reshape(nbins[::-1]), D, rtol=1e-10,
                               atol=1e-12 * D.max())
    vol = np.prod(binsize)
    assert abs(dens.sum() * vol - counts.sum()) <= 1e-10 * counts.sum()


def test_unit_mass_counts_are_integers():
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 10, 5000); y = rng.uniform(0, 10, 5000); z = np.zeros(5000)
    _, counts, _ = O.kde_density(x, y, z, (0, 0, 0), (1, 1, 1), (10, 10, 1))
    assert np.all(counts == np.round(counts)) and counts.sum() == 5000
