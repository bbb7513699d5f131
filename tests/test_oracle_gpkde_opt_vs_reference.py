"""The numpy restatement of the gpkde bandwidth optimisation against the
compiled reference (GridProjectedKDE%ComputeDensityOptimization), loop by
loop through the reference's own export of its optimisation variables
(GridProjectedKDE.F90:5049-5056) and on the final density."""
import numpy as np
import pytest

from oracle import gpkde_opt, refrun

pytestmark = pytest.mark.skipif(not refrun.gpkde_available(),
                                reason="compiled reference gpkde not built")


def cloud(nbins, n=20000, seed=1, weights=False):
    g = np.random.default_rng(seed)
    pts = []
    for d in range(3):
        if nbins[d] > 1:
            c = 0.5 * nbins[d]
            pts.append(np.clip(g.normal(c, nbins[d] / 8.0, n), 0.01, nbins[d] - 0.01))
        else:
            pts.append(g.uniform(0.01, 0.99, n))
    w = g.lognormal(0.0, 0.5, n) if weights else None
    return pts, w


def histogram(pts, nbins, w=None):
    idx = [np.floor(p).astype(np.int64) for p in pts]
    lin = idx[0] + nbins[0] * (idx[1] + nbins[1] * idx[2])
    return np.bincount(lin, weights=w, minlength=int(np.prod(nbins))).astype(float)


def compare_tables(mine, ref, rtol):
    assert mine.shape == ref.shape
    np.testing.assert_array_equal(mine[:, :3], ref[:, :3])
    # the export prints 10 significant digits
    np.testing.assert_allclose(mine[:, 3:], ref[:, 3:], rtol=rtol, atol=1e-300)


CASES = {
    "2d_xy": dict(nbins=(60, 40, 1), loops=3),
    "2d_xz": dict(nbins=(50, 1, 30), loops=3),
    "2d_yz": dict(nbins=(1, 36, 44), loops=2),
    "3d": dict(nbins=(24, 20, 16), loops=2),
    "3d_iso": dict(nbins=(24, 20, 16), loops=2, isotropic=True),
    "2d_iso": dict(nbins=(60, 40, 1), loops=2, isotropic=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_loop_variables(name):
    cfg = CASES[name]
    nb = cfg["nbins"]
    iso = cfg.get("isotropic", False)
    pts, _ = cloud(nb, n=8000 if nb[2] > 1 and nb[0] > 1 else 20000)
    dref, _, h, tabs = refrun.gpkde(*pts, (0, 0, 0), (1, 1, 1), nb,
                                    n_opt_loops=cfg["loops"], isotropic=iso,
                                    unit_volume=False, skip_error=True, export_vars=True)
    counts = histogram(pts, nb)
    dens, loops, mine = gpkde_opt.optimize(counts, nb, (1, 1, 1), h, counts.sum(),
                                           cfg["loops"], isotropic=iso,
                                           skip_error=True, export=True)
    assert loops == cfg["loops"] and len(tabs) == cfg["loops"]
    for a, b in zip(mine, tabs):
        compare_tables(a, b, 5e-8)
    np.testing.assert_allclose(dens, dref, rtol=0, atol=1e-12 * dref.max())


def test_converged_run_and_bins_not_unit():
    """Convergence tests on (no skipped errors), bin sizes other than one."""
    nb = (48, 36, 1)
    bs = (2.5, 0.5, 1.0)
    pts, _ = cloud(nb, n=30000, seed=3)
    pts = [pts[0] * bs[0], pts[1] * bs[1], pts[2]]
    dref, _, h = refrun.gpkde(*pts, (0, 0, 0), bs, nb, n_opt_loops=10,
                              unit_volume=False)
    counts = histogram([pts[0] / bs[0], pts[1] / bs[1], pts[2]], nb)
    dens, loops = gpkde_opt.optimize(counts, nb, bs, h, counts.sum(), 10)
    assert 1 <= loops <= 10
    np.testing.assert_allclose(dens, dref, rtol=0, atol=1e-12 * dref.max())


def test_weighted():
    """Kish effective mass: the loop runs on mass sums / effective mass and
    the density is scaled back by it (Histogram.F90:326-343,
    GridProjectedKDE.F90:4712-4720)."""
    nb = (60, 40, 1)
    pts, w = cloud(nb, n=20000, seed=5, weights=True)
    dref, _, h = refrun.gpkde(*pts, (0, 0, 0), (1, 1, 1), nb, weights=w,
                              n_opt_loops=3, unit_volume=False, skip_error=True)
    eff = w.sum() / (w.sum() ** 2 / (w * w).sum())
    counts = histogram(pts, nb, w) / eff
    dens, loops = gpkde_opt.optimize(counts, nb, (1, 1, 1), h, float(w.size), 3,
                                     skip_error=True)
    np.testing.assert_allclose(dens * eff, dref, rtol=0, atol=1e-12 * dref.max())


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_effective_weight_formats(fmt):
    """Effective weight formats (Histogram.F90:326-343, 352-490): 1 the
    average mass; 2 point counts drive the loop and the mass sums the last
    scatter; 3 per-bin Kish counts (sum w)^2 / sum w^2 with the same last
    scatter (GridProjectedKDE.F90:4712-4760)."""
    nb = (60, 40, 1)
    pts, w = cloud(nb, n=20000, seed=6, weights=True)
    dref, _, h = refrun.gpkde(*pts, (0, 0, 0), (1, 1, 1), nb, weights=w,
                              n_opt_loops=3, unit_volume=False, skip_error=True,
                              weight_format=fmt)
    wsum = histogram(pts, nb, w)
    final, scale = None, 1.0
    if fmt == 1:
        scale = w.sum() / w.size
        counts = wsum / scale
    elif fmt == 2:
        counts, final = histogram(pts, nb), wsum
    else:
        sq = histogram(pts, nb, w * w)
        counts = np.where(sq != 0, wsum ** 2 / np.where(sq != 0, sq, 1.0), 0.0)
        final = wsum
    dens, _ = gpkde_opt.optimize(counts, nb, (1, 1, 1), h, float(w.size), 3,
                                 skip_error=True, final_weights=final)
    np.testing.assert_allclose(dens * scale, dref, rtol=0, atol=1e-12 * dref.max())


@pytest.mark.parametrize("bound,hod", [(1, (0.5, 0.0, 6.0)), (1, (0.9, 0.0, 3.0)),
                                       (2, (0.0, 0.0, 0.0))])
@pytest.mark.parametrize("nb", [(60, 40, 1), (24, 20, 16)])
def test_kernel_size_bounds(bound, hod, nb):
    """boundKernelSizeFormat 1 bounds the kernels by the given h/delta
    range, 2 leaves them unbounded (GridProjectedKDE.F90:788-858, 4863,
    5387, 5551, 5702).  A lower bound with 3 h/delta an integer (h/delta =
    1) is avoided: kernels clamped there sit on the jump of the half-width
    ceiling(3 h/delta), where a last-bit difference of h, from another
    summation order, changes the footprint."""
    pts, _ = cloud(nb, n=8000 if nb[2] > 1 else 20000, seed=8)
    dref, _, h, tabs = refrun.gpkde(*pts, (0, 0, 0), (1, 1, 1), nb, n_opt_loops=2,
                                    unit_volume=False, skip_error=True, bound=bound,
                                    h_over_delta=hod, export_vars=True)
    counts = histogram(pts, nb)
    dens, _, mine = gpkde_opt.optimize(counts, nb, (1, 1, 1), h, counts.sum(), 2,
                                       skip_error=True, export=True, bound=bound,
                                       min_h_over_delta=hod[0], max_h_over_delta=hod[2])
    for a, b in zip(mine, tabs):
        compare_tables(a, b, 5e-8)
    np.testing.assert_allclose(dens, dref, rtol=0, atol=1e-12 * dref.max())


DB_CASES = {
    "2d_xy": ((60, 40, 1), (1.0, 1.0, 1.0)),
    "2d_xz": ((50, 1, 30), (1.0, 1.0, 1.0)),
    "2d_yz": ((1, 36, 44), (1.0, 1.0, 1.0)),
    "2d_xy_aniso": ((60, 40, 1), (2.0, 0.5, 1.0)),
    "3d": ((24, 20, 16), (1.0, 1.0, 1.0)),
    "3d_aniso": ((24, 20, 16), (1.0, 0.5, 2.0)),
}


@pytest.mark.parametrize("loops", [0, 2])
@pytest.mark.parametrize("name", list(DB_CASES))
def test_kernel_database(name, loops):
    """Kernels from the log-spaced database (InitializeKernelDatabaseFlat,
    GridProjectedKDE.F90:3138-3526; lookups :5727-6221), with the lookups'
    quirks: the sigma entry taken without its transpose (larger index in the
    first dimension), and the x-y transpose of an x-z / y-z density lookup
    that leaves the first active dimension one bin wide."""
    nb, bs = DB_CASES[name]
    hod = (0.3, 0.3, 5.0)
    pts, _ = cloud(nb, n=8000 if nb[2] > 1 and nb[0] > 1 else 20000, seed=2)
    pts = [p * b for p, b in zip(pts, bs)]
    res = refrun.gpkde(*pts, (0, 0, 0), bs, nb, n_opt_loops=loops, database=True,
                       h_over_delta=hod, unit_volume=False, skip_error=True,
                       export_vars=loops > 0)
    counts = histogram([p / b for p, b in zip(pts, bs)], nb)
    mine = gpkde_opt.optimize(counts, nb, bs, res[2], counts.sum(), loops,
                              skip_error=True, export=loops > 0, database=hod)
    if loops:
        for a, b in zip(mine[2], res[3]):
            compare_tables(a, b, 5e-8)
    np.testing.assert_allclose(mine[0], res[0], rtol=0, atol=1e-12 * res[0].max())
    brute = refrun.gpkde(*pts, (0, 0, 0), bs, nb, n_opt_loops=loops, unit_volume=False,
                         skip_error=True)[0]
    # the database is in effect: its grid bandwidths change the density
    assert np.abs(brute - res[0]).max() > 1e-6 * res[0].max()
