"""gpkde bandwidth optimisation restated in numpy -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this.
It restates GridProjectedKDE%ComputeDensityOptimization
(src/lib/gpkde/GridProjectedKDE.F90:4396-5160) with brute-force kernels
(prSetKernelBrute, prSetKernelSigmaBrute, prSetKernelSD2DBrute,
prSetKernelSDBrute, :6222-6463) and the three kernel size bound formats,
one active bin at a time in Python loops: for grids of a few thousand
active bins only.  The effective-weight formats enter through the counts
and final weights the caller passes.

Every stage is pinned against the per-loop variables that the compiled
reference exports (exportOptimizationVariables, :5049-5056), see
tests/test_oracle_gpkde_opt_vs_reference.py.
"""
from __future__ import annotations

from math import ceil, floor, log, log10, pi, sqrt

import numpy as np
from scipy.special import erf

RANGE_W = 3          # defaultKernelRange, GridProjectedKDE.F90:21
RANGE_SD = 4         # defaultKernelSDRange
MAX_SIGMA_GROWTH = 1.5
ISO_THRESHOLD = 0.85
ISO_FACTOR = 0.5
INITIAL_SIGMA_FACTOR = 2.0
NLOOPS_ENABLE_AVG_ERROR = 7
_SQ2 = sqrt(2.0)


def _half_width(rng, h, b):
    # matrixPositiveShape = ceiling(matrixRange*bandwidth/binSize),
    # KernelMultiGaussian.f90:569-573
    return 0 if b == 0 else int(ceil(rng * h / b))


def _w_factor(m, hd):
    # erf difference of the W kernel, KernelMultiGaussian.f90:720-790
    i = np.abs(np.arange(-m, m + 1)).astype(float)
    s = hd * _SQ2
    return erf((i + 0.5) / s) - erf((i - 0.5) / s)


def _v_factor(m, hd):
    # first difference of x exp(-x^2/(2 h^2)), the derivative dimension of
    # the V kernels, KernelMultiGaussian.f90:846-849
    i = np.abs(np.arange(-m, m + 1)).astype(float)
    a, b, d = i + 0.5, i - 0.5, 2.0 * hd * hd
    return a * np.exp(-a * a / d) - b * np.exp(-b * b / d)


def _fold(k, m, c, n):
    """Reflection of the kernel entries that fall outside a dimension of n
    bins back into it, for a bin at 1-based index c
    (prKernelReflection, KernelMultiGaussian.f90:303-390)."""
    k = k.copy()
    L = 2 * m + 1
    if c - m < 1:
        b1 = 1 - c + m
        for t in range(1, b1 + 1):
            dst, src = b1 + t, b1 + 1 - t
            if dst <= L:
                k[dst - 1] += k[src - 1]
        k[:min(b1, L)] = 0.0
    if c + m > n:
        b2 = L - (c + m - n) + 1
        lenb = L - b2 + 1
        for t in range(lenb):
            dst, src = b2 - lenb + t, L - t
            if dst >= 1:
                k[dst - 1] += k[src - 1]
        k[b2 - 1:] = 0.0
    return k


class Grid:
    def __init__(self, nbins, bin_size):
        self.nb = [int(v) for v in nbins]
        self.bin = [float(v) for v in bin_size]
        self.active = [v > 1 for v in self.nb]
        self.ndim = sum(self.active)
        self.bin_volume = float(np.prod([b for b in self.bin if b > 0]))


class KernelDatabase:
    """The log-spaced h/delta grid of the kernel database and the kernel each
    lookup returns (InitializeKernelDatabaseFlat, GridProjectedKDE.F90:
    3138-3526; lookups :5727-6221; prGenerateLogSpaceData :6467-6500).
    Every database kernel is a brute-force kernel at a grid bandwidth, so a
    lookup is restated as the bandwidths it stands for."""

    def __init__(self, G, min_hod, delta_hod, max_hod):
        self.G = G
        self.min = float(min_hod)
        n = int(ceil(log10(max_hod / min_hod) / log10(1.0 + delta_hod))) + 1
        lo, hi = log10(min_hod), log10(max_hod)
        step = (hi - lo) / (n - 1)
        acc = 0.0
        self.hod = []
        for _ in range(n):
            self.hod.append(10.0 ** (lo + acc))
            acc = acc + step
        self.n = n
        # the spacing used by the lookups is overwritten with the natural
        # log of the ratio of the first two entries (:3207)
        self.delta = log(self.hod[1] / self.hod[0])
        self.dims = [d for d in range(3) if G.active[d]]

    def index(self, h, d):
        # 1-based, clamped (:5740-5750)
        if h <= 0:
            return 1
        return min(max(int(floor(log(h / self.G.bin[d] / self.min) / self.delta)) + 1, 1),
                   self.n)

    def w(self, h):
        """Density kernel (prSetKernelFromDatabase :5965-6003): bandwidths
        and the dimension collapsed to its centre weight (-1 none).  A 2-D
        lookup with the first index below the second takes the transposed
        entry; the x-y transpose restores the bandwidths in the x-y plane
        and in 3-D, while in the x-z and y-z planes it leaves the first
        active dimension one bin wide (KernelMultiGaussian.f90:393-484)."""
        G = self.G
        idx = [self.index(h[d], d) if d in self.dims else 1 for d in range(3)]
        out = [0.0, 0.0, 0.0]
        if G.ndim == 3:
            for d in range(3):
                out[d] = self.hod[idx[d] - 1] * G.bin[d]
            return out, -1
        a, b = self.dims
        if idx[a] >= idx[b] or (a, b) == (0, 1):
            out[a] = self.hod[idx[a] - 1] * G.bin[a]
            out[b] = self.hod[idx[b] - 1] * G.bin[b]
            return out, -1
        out[a] = self.hod[idx[b] - 1] * G.bin[a]
        out[b] = self.hod[idx[a] - 1] * G.bin[b]
        return out, a

    def sigma(self, s):
        """Sigma kernel (prSetKernelSigmaFromDatabase :6006-6043): the flat
        entry of the per-dimension indices without the transpose, so the
        first two dimensions take the larger and the smaller index."""
        G = self.G
        idx = [self.index(s, d) if d in self.dims else 1 for d in range(3)]
        out = [0.0, 0.0, 0.0]
        if G.ndim == 3:
            hi, lo = max(idx[0], idx[1]), min(idx[0], idx[1])
            out = [self.hod[hi - 1] * G.bin[0], self.hod[lo - 1] * G.bin[1],
                   self.hod[idx[2] - 1] * G.bin[2]]
            return out
        a, b = self.dims
        hi, lo = max(idx[a], idx[b]), min(idx[a], idx[b])
        out[a] = self.hod[hi - 1] * G.bin[a]
        out[b] = self.hod[lo - 1] * G.bin[b]
        return out

    def sd(self, cb, der):
        """Second-derivative kernel along der (prSetKernelSD2DFromDatabase
        :6081-6124, prSetKernelSDFromDatabase :6184-6221): the grid h/delta
        of the derivative dimension's bandwidth in every active dimension."""
        i = self.index(cb, der) if cb > 0 else 1
        return [self.hod[i - 1] * self.G.bin[d] if self.G.active[d] else 0.0
                for d in range(3)]


def kernel(G, c, h, kind, der=0, rng=RANGE_W, collapse=-1):
    """Grid slices and the bounded kernel matrix of one bin.

    kind 0: W kernel (SetupMatrix of KernelMultiGaussianType) with bandwidth
    h[d] in the active dimensions; kind 1: the same kernel used as sigma
    kernel, which carries the bandwidth in every dimension with a bin size
    (an inactive dimension of one bin keeps the centre weight); kind 2: the
    second-derivative kernel V along `der`."""
    ms = []
    for d in range(3):
        hd = h[d] / G.bin[d] if G.bin[d] != 0 else 0.0
        m = 0
        if (kind != 0 or G.active[d]) and hd > 0:
            m = _half_width(rng, h[d], G.bin[d])
        ms.append(m)
    arrs, sums, sq = [], [], []
    for d in range(3):
        hd = h[d] / G.bin[d] if G.bin[d] != 0 else 0.0
        if hd > 0:
            a = _v_factor(ms[d], hd) if (kind == 2 and d == der) else _w_factor(ms[d], hd)
        else:
            a = np.ones(2 * ms[d] + 1)
        arrs.append(a)
        sums.append(a.sum())
        sq.append((a * a).sum())
    if kind in (0, 1):
        # W kernels are normalised to unit sum (:785-788)
        sc = 1.0
        for d in range(3):
            if G.bin[d] != 0 and h[d] > 0:
                sc /= sums[d]
    else:
        hdv = [h[d] / G.bin[d] if G.bin[d] != 0 else 0.0 for d in range(3)]
        fnd = float(G.ndim)
        C = -1.0 / (2.0 ** (fnd - 0.5) * sqrt(pi) * hdv[der] ** 3)
        sc = 0.0 if C * C * sq[0] * sq[1] * sq[2] <= 0 else C
        # the correction loop returns unconditionally at the derivative
        # dimension (KernelMultiGaussian.f90:887, 993, 1097)
        for d in range(der):
            if hdv[d] > 0:
                sc /= sqrt(hdv[d])
    if h[0] == 0 and h[1] == 0 and h[2] == 0:
        sc = 0.0
    if collapse >= 0:
        # only the centre entry of this dimension remains, the normalisation
        # over the full kernel stays
        arrs[collapse] = arrs[collapse][ms[collapse]:ms[collapse] + 1].copy()
        ms[collapse] = 0
    for d in range(3):
        if G.active[d] and ms[d] > 0:
            arrs[d] = _fold(arrs[d], ms[d], c[d], G.nb[d])
    sl, sub = [], []
    for d in range(3):
        lo, hi = max(c[d] - ms[d], 1), min(c[d] + ms[d], G.nb[d])
        sl.append(slice(lo - 1, hi))
        sub.append(arrs[d][lo - c[d] + ms[d]:hi - c[d] + ms[d] + 1])
    K = sc * sub[0][:, None, None] * sub[1][None, :, None] * sub[2][None, None, :]
    return tuple(sl), K


def optimize(counts, nbins, bin_size, h_init, n_points, n_opt_loops,
             isotropic=False, skip_error=False, tol=0.02, export=False,
             bound=0, min_h_over_delta=0.0, max_h_over_delta=0.0,
             final_weights=None, database=None):
    """Adaptive density of a histogram.

    counts: histogram (x fastest) in the units the loop uses (weighted:
    mass sums / effective mass); h_init: initial (Silverman) smoothing;
    n_points: histogram%nPoints; bound: boundKernelSizeFormat (0 domain,
    1 the h/delta range given, 2 none); final_weights: histogram (x fastest)
    scattered with the final kernels in place of the counts (effective
    weight formats 2 and 3, GridProjectedKDE.F90:4712-4760); database:
    (min, delta, max) h/delta of the kernel database, None for brute-force
    kernels.
    Returns (density with x fastest, before the
    effective-mass, unit-volume and scaling factors, loops run, and with
    export the per-loop tables of the reference's export: rows of ix iy iz,
    density, smoothing(3), shape(3), scale, sigma, curvature bandwidth(3),
    nEstimate, roughness XX YY ZZ, net roughness)."""
    G = Grid(nbins, bin_size)
    nb = G.nb
    nd = G.ndim
    if nd < 2:
        raise ValueError("one-dimensional reconstructions are not restated")
    fnd = float(nd)
    grid_counts = np.asarray(counts, float).reshape(nb[::-1]).transpose(2, 1, 0)
    lin = np.flatnonzero(np.asarray(counts, float).ravel() != 0)
    ids = np.stack([lin % nb[0] + 1, (lin // nb[0]) % nb[1] + 1,
                    lin // (nb[0] * nb[1]) + 1], 1)
    N = len(ids)
    cnt = grid_counts[ids[:, 0] - 1, ids[:, 1] - 1, ids[:, 2] - 1]
    binVol = G.bin_volume
    dims = [d for d in range(3) if G.active[d]]
    # prInitializeModuleConstants, GridProjectedKDE.F90:1582-1611
    supportC = ((fnd + 2.0) * (8.0 * pi) ** (0.5 * fnd)) ** 0.25
    alphaC = (((1.0 + 2.0 ** (0.5 * fnd + 2.0)) / (3.0 * 2.0 ** (4.0 / (fnd + 4.0))))
              ** (1.0 / (fnd + 6.0)) * (fnd + 2.0) ** (1.0 / (fnd + 4.0))
              / (fnd + 4.0) ** (1.0 / (fnd + 6.0)))
    betaC = 2.0 / (fnd + 4.0) / (fnd + 6.0)
    # kernel size bounds (:788-858)
    maxKS = np.zeros(3); maxKSD = np.zeros(3); minKS = np.zeros(3); minKSD = np.zeros(3)
    bounded = bound != 2
    for d in dims:
        if bound == 1:
            maxKS[d] = maxKSD[d] = G.bin[d] * max_h_over_delta
            minKS[d] = minKSD[d] = G.bin[d] * min_h_over_delta
        else:
            maxKS[d] = G.bin[d] * (0.9 * nb[d] - 1) / RANGE_W
            maxKSD[d] = G.bin[d] * (0.9 * nb[d] - 1) / RANGE_SD
            minKS[d] = 1.2 * G.bin[d] / RANGE_W
            minKSD[d] = 1.2 * G.bin[d] / RANGE_SD
    maxKSv = min(maxKS[d] for d in dims)
    minKSv = max(minKS[d] for d in dims)

    def smoothing_scale(hk):
        # prComputeKernelSmoothingScale (:5164-5203)
        s = np.ones(N)
        dc = np.full(N, nd)
        for d in dims:
            nz = hk[:, d] != 0
            s = np.where(nz, s * hk[:, d], s)
            dc = dc - (~nz)
        return np.where(dc != 0, s ** (1.0 / fnd), 0.0)

    db = KernelDatabase(G, *database) if database is not None else None

    def density_pass(hk):
        # (:4623-4668, :4930-4960)
        grid = np.zeros(nb)
        for n in range(N):
            if (hk[n] < 0).any():
                continue
            if db is None:
                sl, K = kernel(G, ids[n], hk[n], 0)
            else:
                hw, col = db.w(hk[n])
                sl, K = kernel(G, ids[n], hw, 0, collapse=col)
            grid[sl] += cnt[n] * K
        grid /= binVol
        return grid, grid[ids[:, 0] - 1, ids[:, 1] - 1, ids[:, 2] - 1].copy()

    def sigma_kernel(n, s_):
        if db is None:
            return kernel(G, ids[n], (s_,) * 3, 1)
        return kernel(G, ids[n], db.sigma(s_), 0)

    def gather(grid, sig, f=None):
        out = np.zeros(N)
        for n in range(N):
            sl, K = sigma_kernel(n, sig[n])
            out[n] = (grid[sl] * K).sum()
        return out

    hk = np.zeros((N, 3))
    for d in dims:
        hk[:, d] = h_init[d]
    scale = smoothing_scale(hk)
    sigma = INITIAL_SIGMA_FACTOR * scale
    shape = np.zeros((N, 3))
    for d in dims:
        shape[:, d] = np.where(scale > 0, hk[:, d] / np.where(scale > 0, scale, 1), 0)
    scale_old = scale.copy()
    grid, density = density_pass(hk)
    nEst = np.zeros(N); net = np.zeros(N); curv = np.zeros((N, 3))
    rough = np.zeros((N, 3))

    def almise():
        em = np.where(scale != 0, (nEst / (np.where(scale != 0, scale, 1) ** fnd
                                          * (4.0 * pi) ** (0.5 * fnd))
                                   + 0.25 * net * scale ** 4.0) / n_points ** 2, 0.0)
        return sqrt((em ** 2).sum() / N)

    def table():
        return np.column_stack([ids, density, hk, shape, scale, sigma, curv, nEst,
                                rough, net])

    tables = []
    almise_old = almise()
    cumsum = cumsum_old = almise_old
    density_old = density.copy()
    loops = 0
    for m in range(1, n_opt_loops + 1):
        loops = m
        # support scale and nEstimate, eq. 23 of Sole-Mari et al. (2019)
        # (:4816-4858)
        n1 = gather(grid, sigma)
        ks = n1 ** 0.5 * scale ** (1.0 + 0.25 * fnd) / (4.0 * density) ** 0.25 * supportC
        if bounded:
            ks = np.where(ks > MAX_SIGMA_GROWTH * sigma, MAX_SIGMA_GROWTH * sigma, ks)
            ks = np.minimum(ks, maxKSv)
            ks = np.where(ks < minKSv, minKSv, ks)
        sigma = ks
        nEst = gather(grid, sigma)
        # curvature bandwidths (prComputeCurvatureBandwidth, :5300-5400)
        curv = np.zeros((N, 3))
        ok = density > 0
        nvpb = np.zeros(N)
        nvpb[ok] = ((sqrt(8.0 * pi) * sigma[ok]) ** fnd * nEst[ok] ** 2 / density[ok]) ** betaC
        for did in dims:
            num = np.ones(3)
            num[did] = 5.0
            s = sum(num[dd] / shape[:, dd] ** 2 for dd in dims)
            with np.errstate(divide="ignore", invalid="ignore"):
                term = ((1.0 / (fnd + 4.0)) / shape[:, did] ** 4 * s) ** (-1.0 / (fnd + 6.0))
            cb = alphaC * nvpb * term * scale
            if bounded:
                cb = np.where(cb > maxKSD[did], maxKSD[did], cb)
                cb = np.where(cb < minKSD[did], minKSD[did], cb)
            curv[:, did] = np.where(ok, cb, 0.0)
        # net roughness (prComputeNetRoughness2D :1905-2289, 3D :2295-2709)
        rough = np.zeros((N, 3))
        ders = dims if nd == 2 else [0, 1, 2]
        cgs = []
        for der in ders:
            cg = np.zeros(nb)
            for n in range(N):
                if nd == 2 and ((curv[n] < 0).any() or np.isnan(curv[n]).any()):
                    continue
                if db is None:
                    h = [curv[n, der] if G.active[d] else 0.0 for d in range(3)]
                else:
                    h = db.sd(curv[n, der], der)
                sl, K = kernel(G, ids[n], h, 2, der, RANGE_SD)
                cg[sl] += cnt[n] * K
            cgs.append(cg / binVol / G.bin[der] ** 2)
        if nd == 2:
            c1, c2 = cgs
            c12 = c1 * c2
            s = np.zeros((N, 3))
            for n in range(N):
                sl, K = sigma_kernel(n, sigma[n])
                s[n] = [(c1[sl] ** 2 * K).sum(), (c2[sl] ** 2 * K).sum(), (c12[sl] * K).sum()]
            rough[:, dims[0]] = s[:, 0]
            rough[:, dims[1]] = s[:, 1]
            if not isotropic:
                net = 2.0 * s[:, 2] + 2.0 * np.sqrt(s[:, 0] * s[:, 1])
            else:
                net = 2.0 * s[:, 2] + s[:, 0] + s[:, 1]
        else:
            # cross products of the squared curvatures (:2464-2469)
            cx, cy, cz = (c * c for c in cgs)
            s = np.zeros((N, 6))
            for n in range(N):
                sl, K = sigma_kernel(n, sigma[n])
                s[n] = [(cx[sl] * K).sum(), (cy[sl] * K).sum(), (cz[sl] * K).sum(),
                        (cx[sl] * cy[sl] * K).sum(), (cx[sl] * cz[sl] * K).sum(),
                        (cy[sl] * cz[sl] * K).sum()]
            rough[:] = s[:, :3]
            rXY, rXZ, rYZ = s[:, 3], s[:, 4], s[:, 5]
            rXX, rYY, rZZ = s[:, 0], s[:, 1], s[:, 2]
            if not isotropic:
                with np.errstate(divide="ignore", invalid="ignore"):
                    v = 3.0 * (rXX * rYY * rZZ) ** (1.0 / 3.0)
                    v = v + np.where(rYZ > 0, 2.0 * rYZ * (rXX ** 2 / rYY / rZZ) ** (1.0 / 6.0), 0.0)
                    v = v + np.where(rXZ > 0, 2.0 * rXZ * (rYY ** 2 / rXX / rZZ) ** (1.0 / 6.0), 0.0)
                    v = v + np.where(rXY > 0, 2.0 * rXY * (rZZ ** 2 / rXX / rYY) ** (1.0 / 6.0), 0.0)
                net = v
            else:
                net = rXX + 2.0 * rXY + 2.0 * rXZ + rYY + 2.0 * rYZ + rZZ
        # optimal smoothing and shape (:5404-5724)
        r11 = rough[:, dims[0]]
        r22 = rough[:, dims[1]]
        if not isotropic:
            minNet = net[net > 0].min() if (net > 0).any() else np.inf
            sumMain = r11 + r22 if nd == 2 else rough.sum(1)
            sumMain = np.where(sumMain == 0, minNet, sumMain)
            with np.errstate(divide="ignore", invalid="ignore"):
                if nd == 2:
                    iso = (r11 / sumMain > ISO_THRESHOLD) | (r22 / sumMain > ISO_THRESHOLD)
                else:
                    iso = ((rough[:, 0] / sumMain > ISO_THRESHOLD)
                           | (rough[:, 1] / sumMain > ISO_THRESHOLD)
                           | (rough[:, 2] / sumMain > ISO_THRESHOLD))
            net = np.where(iso, (1.0 - ISO_FACTOR) * net + ISO_FACTOR * sumMain, net)
        scale = np.zeros(N)
        pos = net > 0
        scale[pos] = (fnd * nEst[pos] / ((4.0 * pi) ** (0.5 * fnd) * net[pos])) ** (1.0 / (fnd + 4.0))
        if (scale > 0).any():
            scale[net == 0] = scale[scale > 0].max()
        shape = np.ones((N, 3))
        if not isotropic:
            for n in range(N):
                if nd == 2:
                    if r11[n] == 0 or r22[n] == 0:
                        continue
                    a, b = dims
                    shape[n, a] = (net[n] / r11[n]) ** 0.125
                    if bounded:
                        shape[n, a] = max(shape[n, a], minKS[a] / scale[n])
                    shape[n, b] = 1.0 / shape[n, a]
                    ns = sqrt(shape[n, a] * shape[n, b])
                    if ns == 0:
                        shape[n] = 1.0
                        continue
                    shape[n, a] /= ns
                    shape[n, b] /= ns
                else:
                    if (rough[n] == 0).any():
                        continue
                    shape[n, 0] = (net[n] / rough[n, 0]) ** (1.0 / 12.0)
                    shape[n, 1] = (net[n] / rough[n, 1]) ** (1.0 / 12.0)
                    if bounded:
                        shape[n, 0] = max(shape[n, 0], minKS[0] / scale[n])
                        shape[n, 1] = max(shape[n, 1], minKS[1] / scale[n])
                    shape[n, 2] = 1.0 / shape[n, 0] / shape[n, 1]
                    # single-precision literal 0.333 in the reference (:5640)
                    ns = (shape[n, 0] * shape[n, 1] * shape[n, 2]) ** float(np.float32(0.333))
                    if ns == 0:
                        shape[n] = 1.0
                        continue
                    shape[n] /= ns
        hk = np.zeros((N, 3))
        update = False
        for d in dims:
            hk[:, d] = shape[:, d] * scale
            if not bounded:
                continue
            if (hk[:, d] > maxKS[d]).any():
                update = True
                hk[:, d] = np.minimum(hk[:, d], maxKS[d])
            if (hk[:, d] < minKS[d]).any():
                update = True
                hk[:, d] = np.maximum(hk[:, d], minKS[d])
        if update:
            scale = smoothing_scale(hk)
        grid, density = density_pass(hk)
        # convergence (:4962-5038)
        mD, mO = density.max(), density_old.max()
        with np.errstate(divide="ignore", invalid="ignore"):
            em = np.where(density_old != 0,
                          np.abs(density / mD - density_old / mO) / (density_old / mO), 0.0)
            err_density = sqrt((em ** 2).sum() / N)
            es = np.where(scale_old != 0, np.abs(scale - scale_old) / scale_old, 0.0)
            err_smoothing = sqrt((es ** 2).sum() / N)
        alm = almise()
        cumsum = cumsum + alm
        if not skip_error:
            if almise_old > 0 and abs(alm - almise_old) / almise_old < tol:
                break
            if err_density < tol or err_smoothing < tol:
                break
            if m > NLOOPS_ENABLE_AVG_ERROR and (
                    abs(cumsum_old / (m - 1) - cumsum / m) / cumsum_old / (m - 1) < tol):
                break
        almise_old, cumsum_old = alm, cumsum
        density_old = density.copy()
        scale_old = scale.copy()
        if export:
            tables.append(table())
    if final_weights is not None:
        fw = np.asarray(final_weights, float).reshape(nb[::-1]).transpose(2, 1, 0)
        cnt = fw[ids[:, 0] - 1, ids[:, 1] - 1, ids[:, 2] - 1]
        grid, density = density_pass(hk)
    out = grid.transpose(2, 1, 0).ravel().copy()
    return (out, loops, tables) if export else (out, loops)
