"""GPU parity of the gpkde bandwidth optimisation (nOptimizationLoops > 0),
its options and the kernel database.

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
    # name: grid, bins (size, count), loops, weighted, isotropic, unit volume,
    # scaling factor, further options
    "2d": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 10, False, False, False, 1.0, {}),
    "2d_weighted": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, True, False, False,
                    1.0, {}),
    "2d_iso": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, False, True, False, 1.0, {}),
    "2d_bins": ((60, 40, 1), ((2.0, 0.5, 1.0), (30, 80, 1)), 10, False, False, True, 0.3,
                {}),
    "3d": ((24, 20, 12), ((1.0, 1.0, 1.0), (24, 20, 12)), 4, False, False, False, 1.0, {}),
    "3d_weighted_iso": ((24, 20, 12), ((1.0, 1.0, 1.0), (24, 20, 12)), 3, True, True,
                        True, 2.0, {}),
    "2d_weight_format1": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 3, True, False,
                          False, 1.0, {"effective_weight_format": 1}),
    "2d_weight_format2": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 3, True, False,
                          False, 1.0, {"effective_weight_format": 2}),
    "3d_weight_format3": ((24, 20, 12), ((1.0, 1.0, 1.0), (24, 20, 12)), 3, True, False,
                          True, 1.0, {"effective_weight_format": 3}),
    "2d_bound1": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, False, False, False,
                  1.0, {"bound_kernel_size": 1, "min_h_over_delta": 0.9,
                        "max_h_over_delta": 3.0}),
    "3d_unbounded": ((24, 20, 12), ((1.0, 1.0, 1.0), (24, 20, 12)), 3, False, False,
                     False, 1.0, {"bound_kernel_size": 2}),
    "2d_smoothing_factor": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, False, False,
                            False, 1.0, {"initial_smoothing": 1, "smoothing_factor": 1.7}),
    "2d_compressed_z": ((60, 40, 1), ((1.0, 1.0, 0.0), (60, 40, 1)), 4, False, False,
                        False, 1.0, {}),
    "2d_database": ((60, 40, 1), ((1.0, 1.0, 1.0), (60, 40, 1)), 4, False, False, False,
                    1.0, {"kernel_database": True, "min_h_over_delta": 0.3,
                          "delta_h_over_delta": 0.3, "max_h_over_delta": 5.0}),
    "2d_database_aniso_bins": ((60, 40, 1), ((2.0, 0.5, 1.0), (30, 80, 1)), 4, True, False,
                               False, 1.0, {"kernel_database": True,
                                            "min_h_over_delta": 0.3,
                                            "delta_h_over_delta": 0.3,
                                            "max_h_over_delta": 6.0}),
    "2d_xz_database": ((50, 1, 30), ((1.0, 1.0, 1.0), (50, 1, 30)), 3, False, False, False,
                       1.0, {"kernel_database": True, "min_h_over_delta": 0.3,
                             "delta_h_over_delta": 0.3, "max_h_over_delta": 5.0}),
    "2d_xz_database_no_loops": ((50, 1, 30), ((1.0, 1.0, 1.0), (50, 1, 30)), 0, False,
                                False, True, 1.0, {"kernel_database": True,
                                                   "min_h_over_delta": 0.3,
                                                   "delta_h_over_delta": 0.3,
                                                   "max_h_over_delta": 5.0}),
    "3d_database": ((24, 20, 12), ((1.0, 0.5, 2.0), (24, 20, 12)), 3, False, False, False,
                    1.0, {"kernel_database": True, "min_h_over_delta": 0.3,
                          "delta_h_over_delta": 0.3, "max_h_over_delta": 5.0}),
}


def _ref_kwargs(kw):
    """The same options in the compiled reference driver's terms."""
    out = {}
    if "effective_weight_format" in kw:
        out["weight_format"] = kw["effective_weight_format"]
    if "bound_kernel_size" in kw:
        out["bound"] = kw["bound_kernel_size"]
    if "bound_kernel_size" in kw or "kernel_database" in kw:
        out["h_over_delta"] = (kw.get("min_h_over_delta", 0.0),
                               kw.get("delta_h_over_delta", 0.0),
                               kw.get("max_h_over_delta", 0.0))
    if kw.get("kernel_database"):
        out["database"] = True
    if "initial_smoothing" in kw:
        out["smoothing_selection"] = kw["initial_smoothing"]
        out["smoothing_factor"] = kw["smoothing_factor"]
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_adaptive_density(name):
    grid, (bs, nb), loops, weighted, iso, uv, sf, kw = CASES[name]
    P = unit_grid(*grid)
    cell, lx, ly, lz = cloud(P, 20000, 11)
    mass = np.random.default_rng(2).lognormal(0.0, 0.4, cell.size) if weighted else None
    origin = (0.0, 0.0, 0.0)
    print(f"gpkde opt {name}: grid {grid} bins {nb} loops {loops} {kw}", flush=True)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        dens, hist = E.kde(origin, bs, nb, mode=1, mass=mass, scaling_factor=sf,
                           unit_volume=uv, n_opt_loops=loops, isotropic=iso, **kw)
        info = E.kde_info()
    finally:
        E.close()
    gx, gy, gz = O.global_xyz(P, cell, lx, ly, lz)
    od, oc, oh = O.kde_density(gx, gy, gz, origin, bs, nb, mass,
                               kw.get("initial_smoothing", 0), kw.get("smoothing_factor", 1.0))
    assert abs(info["bandwidth"] - oh[0]) <= 1e-12 * oh[0]
    fmt = kw.get("effective_weight_format", 0)
    final, eff = None, 1.0
    if mass is None:
        assert np.array_equal(hist, oc)
        counts, npts = oc, oc.sum()
    else:
        wsum = O.histogram(gx, gy, gz, origin, bs, nb, mass)
        np.testing.assert_allclose(hist, wsum, rtol=1e-14, atol=0)
        npts = float(mass.size)
        if fmt == 0:
            eff = mass.sum() / (mass.sum() ** 2 / (mass * mass).sum())
        elif fmt == 1:
            eff = mass.sum() / mass.size
        if fmt in (0, 1):
            counts = wsum / eff
        elif fmt == 2:
            counts, final = O.histogram(gx, gy, gz, origin, bs, nb), wsum
        else:
            sq = O.histogram(gx, gy, gz, origin, bs, nb, mass * mass)
            counts = np.where(sq != 0, wsum ** 2 / np.where(sq != 0, sq, 1.0), 0.0)
            final = wsum
    want, oloops = gpkde_opt.optimize(
        counts, nb, bs, oh, npts, loops, isotropic=iso,
        bound=kw.get("bound_kernel_size", 0),
        min_h_over_delta=kw.get("min_h_over_delta", 0.0),
        max_h_over_delta=kw.get("max_h_over_delta", 0.0), final_weights=final,
        database=((kw["min_h_over_delta"], kw["delta_h_over_delta"], kw["max_h_over_delta"])
                  if kw.get("kernel_database") else None))
    want = want * eff * sf * (np.prod([b for b in bs if b > 0]) if uv else 1.0)
    assert info["loops"] == oloops
    np.testing.assert_allclose(dens, want, rtol=0, atol=1e-9 * want.max())
    if refrun.gpkde_available():
        dr, _, _ = refrun.gpkde(gx, gy, gz, origin, bs, nb, weights=mass, n_opt_loops=loops,
                                isotropic=iso, unit_volume=uv, scaling_factor=sf,
                                **_ref_kwargs(kw))
        np.testing.assert_allclose(dens, dr, rtol=0, atol=1e-9 * dr.max())


def test_single_bin_dimension_drops_points_outside():
    """A one-bin dimension with a bin size drops the points outside it, as
    the reference's histogram does (Histogram.F90:306-317); bin size 0
    keeps them."""
    P = unit_grid(60, 40, 1)
    cell, lx, ly, lz = cloud(P, 6000, 12)
    E = Engine(P)
    try:
        E.load_particles(cell, lx, ly, lz)
        d1, h1 = E.kde((0.0, 0.0, 0.5), (1.0, 1.0, 1.0), (60, 40, 1), mode=1)
        d0, h0 = E.kde((0.0, 0.0, 0.5), (1.0, 1.0, 0.0), (60, 40, 1), mode=1)
    finally:
        E.close()
    gx, gy, gz = O.global_xyz(P, cell, lx, ly, lz)
    inside = (gz >= 0.5) & (gz < 1.5)
    assert h1.sum() == inside.sum() < h0.sum() == cell.size
    for bs, d, h in (((1.0, 1.0, 1.0), d1, h1), ((1.0, 1.0, 0.0), d0, h0)):
        od, oc, _ = O.kde_density(gx, gy, gz, (0.0, 0.0, 0.5), bs, (60, 40, 1))
        assert np.array_equal(h, oc)
        np.testing.assert_allclose(d, od, rtol=1e-10, atol=1e-13 * od.max())


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
