"""Synthetic MODPATH-RW problems: structured grids, face flows, particle clouds.

The reference reads MODFLOW outputs (FlowModelData.f90:217-622); here the
same arrays are synthesised so that the hot path can run without MODFLOW:

* face flows follow the MODFLOW-2005 conventions that
  ModpathCellData.f90:1315-1351 consumes: ``frf`` = flow from column j to
  j+1, ``fff`` = flow from row i to i+1, ``flf`` = flow from layer k to k+1,
  and ``bflow`` holds 6 boundary flows per cell (positive into the cell);
* cell numbers are 1-based, (k-1)*nrow*ncol + (i-1)*ncol + j, arrays are
  ordered the same way (layer slowest, column fastest);
* the heterogeneous field is the discrete curl of a vector potential built
  from seeded lognormal fields, so every cell balances to rounding.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np


@dataclasses.dataclass
class Problem:
    ncol: int
    nrow: int
    nlay: int
    delr: np.ndarray
    delc: np.ndarray
    top: np.ndarray
    bot: np.ndarray
    frf: np.ndarray
    fff: np.ndarray
    flf: np.ndarray
    porosity: np.ndarray
    bflow: Optional[np.ndarray] = None
    retardation: Optional[np.ndarray] = None
    ibound: Optional[np.ndarray] = None
    icbound: Optional[np.ndarray] = None
    default_icbound: int = 0
    disp_model: int = 0
    aL: np.ndarray = dataclasses.field(default_factory=lambda: np.array([0.1]))
    aT: np.ndarray = dataclasses.field(default_factory=lambda: np.array([0.01]))
    aLV: Optional[np.ndarray] = None
    aTV: Optional[np.ndarray] = None
    dmeff: np.ndarray = dataclasses.field(default_factory=lambda: np.array([1e-4]))
    advection_kind: int = 2      # 1 exponential, 2 eulerian
    time_step_kind: int = 4      # 1 ADV, 2 DISP, 3 MIN_ADV_DISP, 4 FIXED
    ts_params: tuple = (0.25, 0.0)
    dim_mask: tuple = (1, 1, 1)
    backward: int = 0
    # cell extras (rwpt_set_cell_extras; None: not given): heads and layer
    # types of convertible layers, zones, weak source and sink flows
    head: Optional[np.ndarray] = None
    layer_type: Optional[np.ndarray] = None
    zone: Optional[np.ndarray] = None
    source_flow: Optional[np.ndarray] = None
    sink_flow: Optional[np.ndarray] = None
    # stopping options (ParticleTrackingOptions.f90:20-31) and
    # FlowModelData%SteadyState
    stop_zone: int = 0
    stop_weak_sinks: int = 0
    stop_weak_sources: int = 0
    steady_state: int = 1
    # Grid%SaturatedTop of the time step (the top of ConvertToModelXYZ),
    # None: the cell tops
    saturated_top: Optional[np.ndarray] = None

    @property
    def has_extras(self) -> bool:
        return any(a is not None for a in (self.head, self.layer_type, self.zone,
                                           self.source_flow, self.sink_flow))

    @property
    def ncell(self) -> int:
        return self.ncol * self.nrow * self.nlay

    @property
    def disp_uniform(self) -> int:
        return int(self.aL.size == 1)

    @property
    def uniform_porosity(self) -> int:
        """isUniformPorosity, ModpathBasicData.f90:208."""
        return int(bool(np.all(self.porosity == self.porosity[0])))

    def cell_number(self, k, i, j):
        """1-based MODFLOW cell number from 1-based layer, row, column."""
        return (np.asarray(k) - 1) * self.nrow * self.ncol + \
            (np.asarray(i) - 1) * self.ncol + np.asarray(j)


def _uniform_grid(ncol, nrow, nlay, d=1.0):
    delr = np.full(ncol, d)
    delc = np.full(nrow, d)
    ncell = ncol * nrow * nlay
    k = np.repeat(np.arange(nlay), nrow * ncol)
    top = (nlay - k) * d
    bot = top - d
    return delr, delc, top.astype(np.float64), bot.astype(np.float64)


def uniform_flow(ncol=100, nrow=50, nlay=1, qx=1.0, porosity=0.3, d=1.0):
    """Uniform flow q_x on every x-face, including the two boundary faces."""
    delr, delc, top, bot = _uniform_grid(ncol, nrow, nlay, d)
    ncell = ncol * nrow * nlay
    area = d * d
    frf = np.zeros(ncell)
    col = np.tile(np.arange(1, ncol + 1), nrow * nlay)
    frf[col < ncol] = qx * area
    fff = np.zeros(ncell)
    flf = np.zeros(ncell)
    bflow = np.zeros(6 * ncell)
    first = np.nonzero(col == 1)[0]
    last = np.nonzero(col == ncol)[0]
    bflow[6 * first + 0] = qx * area       # into the cell through face 1
    bflow[6 * last + 1] = -qx * area       # out of the cell through face 2
    return Problem(ncol=ncol, nrow=nrow, nlay=nlay, delr=delr, delc=delc,
                   top=top, bot=bot, frf=frf, fff=fff, flf=flf,
                   porosity=np.full(ncell, porosity), bflow=bflow)


def config0(**kw):
    """100x50x1, q_x=1, phi=0.3, aL=0.1, aT=0.01, Dm=1e-4, 2-D, FIXED 0.25 d."""
    P = uniform_flow(100, 50, 1, qx=1.0, porosity=0.3)
    P.aL = np.array([0.1])
    P.aT = np.array([0.01])
    P.dmeff = np.array([1e-4])
    P.dim_mask = (1, 1, 0)
    P.ts_params = (0.25, 0.0)
    P.time_step_kind = 4
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def gaussian_field(shape, corr_len, sigma, rng):
    """Stationary Gaussian field with exponential covariance
    sigma^2 exp(-r/corr_len) by circulant embedding on a padded periodic grid
    (FFT); the clipped negative eigenvalues make it approximate at the edges
    of the padding only."""
    pad = [int(2 ** np.ceil(np.log2(n + 3 * corr_len))) if n > 1 else 1
           for n in shape]
    grids = np.meshgrid(*[np.minimum(np.arange(p), p - np.arange(p)).astype(float)
                          for p in pad], indexing="ij")
    r = np.sqrt(sum(g * g for g in grids))
    cov = sigma ** 2 * np.exp(-r / corr_len)
    lam = np.real(np.fft.fftn(cov))
    lam[lam < 0] = 0.0
    noise = rng.standard_normal(pad) + 1j * rng.standard_normal(pad)
    f = np.fft.ifftn(np.sqrt(lam) * noise) * np.sqrt(np.prod(pad))
    f = np.real(f)[tuple(slice(0, n) for n in shape)]
    return f


def heterogeneous_flow(ncol=200, nrow=200, nlay=20, qmean=0.5, sigma_lnk=1.0,
                       corr_len=10.0, seed=20240601, pert=0.5,
                       porosity_range=(0.20, 0.30), d=1.0):
    """Divergence-free face flows = mean q_x plus the discrete curl of a
    seeded vector potential A whose three components are lognormal fields
    K_c/mean(K_c) - 1 (sigma_lnK, exponential covariance, corr_len cells).
    Tangential A is zero on the domain boundary, so only the mean flow
    crosses the two x-boundaries.  The curl is rescaled so that the standard
    deviation of the perturbation x-flux equals ``pert * qmean``."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = ncol, nrow, nlay
    # edge-centred potential components (local axes: x=column, y=-row, z=up)
    def lognormal(shape):
        y = gaussian_field(shape, corr_len, sigma_lnk, rng)
        k = np.exp(y)
        return k / k.mean() - 1.0
    Ax = lognormal((nx, ny + 1, nz + 1))
    Ay = lognormal((nx + 1, ny, nz + 1))
    Az = lognormal((nx + 1, ny + 1, nz))
    Ax[:, 0, :] = Ax[:, -1, :] = 0.0
    Ax[:, :, 0] = Ax[:, :, -1] = 0.0
    Ay[0, :, :] = Ay[-1, :, :] = 0.0
    Ay[:, :, 0] = Ay[:, :, -1] = 0.0
    Az[0, :, :] = Az[-1, :, :] = 0.0
    Az[:, 0, :] = Az[:, -1, :] = 0.0
    # face fluxes (volumetric, d = 1 m cells) of curl A
    Qx = (Az[:, 1:, :] - Az[:, :-1, :]) * d - (Ay[:, :, 1:] - Ay[:, :, :-1]) * d
    Qy = (Ax[:, :, 1:] - Ax[:, :, :-1]) * d - (Az[1:, :, :] - Az[:-1, :, :]) * d
    Qz = (Ay[1:, :, :] - Ay[:-1, :, :]) * d - (Ax[:, 1:, :] - Ax[:, :-1, :]) * d
    s = Qx[1:-1].std()
    scale = pert * qmean * d * d / s if s > 0 else 0.0
    Qx *= scale
    Qy *= scale
    Qz *= scale
    Qx += qmean * d * d
    delr, delc, top, bot = _uniform_grid(ncol, nrow, nlay, d)
    ncell = ncol * nrow * nlay
    # MODFLOW arrays in (layer, row, column) order
    frf = np.zeros((nlay, nrow, ncol))
    fff = np.zeros((nlay, nrow, ncol))
    flf = np.zeros((nlay, nrow, ncol))
    # local y index iy = nrow-1-i0, z index iz = nlay-1-k0 (0-based row/layer)
    qx_cell_right = Qx[1:, :, :]                   # (nx, ny, nz) right faces
    frf[:] = np.transpose(qx_cell_right[:, ::-1, ::-1], (2, 1, 0))
    frf[:, :, -1] = 0.0
    # front face of row i0 is the lower-y face of local iy: node iy
    qy_front = Qy[:, :-1, :]                        # lower-y face of each cell
    fff[:] = -np.transpose(qy_front[:, ::-1, ::-1], (2, 1, 0))
    fff[:, -1, :] = 0.0
    qz_bottom = Qz[:, :, :-1]                       # lower-z face of each cell
    flf[:] = -np.transpose(qz_bottom[:, ::-1, ::-1], (2, 1, 0))
    flf[-1, :, :] = 0.0
    bflow = np.zeros((nlay, nrow, ncol, 6))
    bflow[:, :, 0, 0] = qmean * d * d
    bflow[:, :, -1, 1] = -qmean * d * d
    lo, hi = porosity_range
    phi = rng.uniform(lo, hi, size=ncell)
    return Problem(ncol=ncol, nrow=nrow, nlay=nlay, delr=delr, delc=delc,
                   top=top, bot=bot, frf=frf.ravel(), fff=fff.ravel(),
                   flf=flf.ravel(), porosity=phi, bflow=bflow.ravel())


def config1(ncol=200, nrow=200, nlay=20, seed=20240601, **kw):
    """Heterogeneous 200x200x20 field, aL=0.5, aT=0.05, Dm=1e-4, FIXED 0.5 d."""
    P = heterogeneous_flow(ncol, nrow, nlay, seed=seed)
    P.aL = np.array([0.5])
    P.aT = np.array([0.05])
    P.dmeff = np.array([1e-4])
    P.dim_mask = (1, 1, 1)
    P.ts_params = (0.5, 0.0)
    P.time_step_kind = 4
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def release_in_cell(P: Problem, n, k, i, j, seed=1, lz=None):
    """n particles uniformly inside cell (k, i, j) (1-based)."""
    rng = np.random.default_rng(seed)
    cell = np.full(n, int(P.cell_number(k, i, j)), dtype=np.int32)
    lx = rng.uniform(0.0, 1.0, n)
    ly = rng.uniform(0.0, 1.0, n)
    lzv = rng.uniform(0.0, 1.0, n) if lz is None else np.full(n, float(lz))
    return cell, lx, ly, lzv


def release_in_block(P: Problem, n, cols, rows, lays, seed=1):
    """n particles uniformly inside the block of cells with 1-based inclusive
    column, row and layer ranges (equal-size cells)."""
    rng = np.random.default_rng(seed)
    j = rng.integers(cols[0], cols[1] + 1, n)
    i = rng.integers(rows[0], rows[1] + 1, n)
    k = rng.integers(lays[0], lays[1] + 1, n)
    cell = P.cell_number(k, i, j).astype(np.int32)
    return cell, rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)


BASE_SEED = (0x1DADBEEFBAADD0D0, 0x5BADD0D0DEADBEEF)
"""baseseed of pr_InitializeRandomWalk, TrackSubCell.F90:553-555."""


@dataclasses.dataclass
class Ensemble(Problem):
    """Members sharing one grid (ncol, nrow, nlay, delr, delc, top, bot of the
    first member), with their per-cell arrays concatenated member-major as
    rwpt_create_ensemble expects them; dispersion is one value per member
    when every member's is uniform."""
    members: int = 1

    @property
    def ncell(self) -> int:
        return self.ncol * self.nrow * self.nlay * self.members

    @property
    def ncell_member(self) -> int:
        return self.ncol * self.nrow * self.nlay

    @property
    def disp_uniform(self) -> int:
        return 2 if self.aL.size == self.members and self._per_member else 0

    _per_member: bool = True

    def global_cell(self, member, cell):
        """Member-major cell number of a member's 1-based cell."""
        return np.asarray(member) * self.ncell_member + np.asarray(cell)


def stack(problems):
    """Ensemble of member problems on the same grid."""
    p0 = problems[0]
    for q in problems[1:]:
        if (q.ncol, q.nrow, q.nlay) != (p0.ncol, p0.nrow, p0.nlay):
            raise ValueError("ensemble members must share the grid")
        for name in ("delr", "delc", "top", "bot"):
            if not np.array_equal(getattr(q, name), getattr(p0, name)):
                raise ValueError(f"ensemble members must share {name}")
        for name in ("default_icbound", "disp_model", "advection_kind",
                     "time_step_kind", "ts_params", "dim_mask", "backward"):
            if getattr(q, name) != getattr(p0, name):
                raise ValueError(f"ensemble members must share {name}")
    n = p0.ncell

    def cat(name, size, fill):
        parts = []
        for q in problems:
            a = getattr(q, name)
            parts.append(np.full(size, fill) if a is None else np.asarray(a))
        return np.concatenate(parts)

    per_member = all(q.aL.size == 1 for q in problems)

    def disp(name, fill):
        if per_member:
            return np.array([fill if getattr(q, name) is None else
                             float(np.asarray(getattr(q, name)).ravel()[0])
                             for q in problems])
        return np.concatenate([np.full(n, fill) if getattr(q, name) is None else
                               np.broadcast_to(np.asarray(getattr(q, name), float), (n,))
                               for q in problems])

    E = Ensemble(
        ncol=p0.ncol, nrow=p0.nrow, nlay=p0.nlay, delr=p0.delr, delc=p0.delc,
        top=p0.top, bot=p0.bot, frf=cat("frf", n, 0.0), fff=cat("fff", n, 0.0),
        flf=cat("flf", n, 0.0), porosity=cat("porosity", n, 0.0),
        bflow=cat("bflow", 6 * n, 0.0), retardation=cat("retardation", n, 1.0),
        ibound=cat("ibound", n, 1).astype(np.int32),
        icbound=cat("icbound", n, 0).astype(np.int32),
        default_icbound=p0.default_icbound, disp_model=p0.disp_model,
        aL=disp("aL", 0.0), aT=disp("aT", 0.0),
        aLV=disp("aLV", 0.0) if p0.disp_model == 1 else None,
        aTV=disp("aTV", 0.0) if p0.disp_model == 1 else None,
        dmeff=disp("dmeff", 0.0), advection_kind=p0.advection_kind,
        time_step_kind=p0.time_step_kind, ts_params=p0.ts_params,
        dim_mask=p0.dim_mask, backward=p0.backward, members=len(problems),
        saturated_top=None if all(q.saturated_top is None for q in problems) else
        np.concatenate([q.top if q.saturated_top is None else q.saturated_top
                        for q in problems]))
    E._per_member = per_member
    return E


def config3(members=256, seed=20240602, n_min=10_000, n_max=1_000_000):
    """Ensemble of `members` on one seeded 100x100x10 heterogeneous field
    (the generator of config1): per member aL ~ U[0.05, 1] m,
    aT = aL * U[0.05, 0.2], a uniform porosity ~ U[0.2, 0.35] and a particle
    count log-uniform in [n_min, n_max].  Returns (ensemble, counts)."""
    base = heterogeneous_flow(100, 100, 10, seed=seed)
    rng = np.random.default_rng(seed + 1)
    aL = rng.uniform(0.05, 1.0, members)
    aT = aL * rng.uniform(0.05, 0.2, members)
    phi = rng.uniform(0.2, 0.35, members)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), members)).astype(np.int64)
    n = base.ncell
    E = Ensemble(ncol=base.ncol, nrow=base.nrow, nlay=base.nlay, delr=base.delr,
                 delc=base.delc, top=base.top, bot=base.bot,
                 frf=np.tile(base.frf, members), fff=np.tile(base.fff, members),
                 flf=np.tile(base.flf, members), porosity=np.repeat(phi, n),
                 bflow=np.tile(base.bflow, members), aL=aL, aT=aT,
                 dmeff=np.full(members, 1e-4), ts_params=(0.5, 0.0),
                 members=members)
    return E, counts


def member(E: Ensemble, m: int) -> Problem:
    """Member m of an ensemble as a problem of its own (the same member run
    alone, with stream_offset = its first particle's index, reproduces its
    particles in the ensemble)."""
    n = E.ncell_member
    cells = slice(m * n, (m + 1) * n)

    def cut(a, per=1):
        if a is None:
            return None
        a = np.asarray(a)
        return a[per * m * n:per * (m + 1) * n].copy()

    def disp(a):
        if a is None:
            return None
        a = np.asarray(a)
        if a.size == E.members and E._per_member:
            return a[m:m + 1].copy()
        return a[cells].copy()

    return Problem(ncol=E.ncol, nrow=E.nrow, nlay=E.nlay, delr=E.delr, delc=E.delc,
                   top=np.asarray(E.top)[:n], bot=np.asarray(E.bot)[:n],
                   frf=cut(E.frf), fff=cut(E.fff), flf=cut(E.flf), porosity=cut(E.porosity),
                   bflow=cut(E.bflow, 6), retardation=cut(E.retardation),
                   ibound=cut(E.ibound), icbound=cut(E.icbound),
                   default_icbound=E.default_icbound, disp_model=E.disp_model,
                   aL=disp(E.aL), aT=disp(E.aT), aLV=disp(E.aLV), aTV=disp(E.aTV),
                   dmeff=disp(E.dmeff), advection_kind=E.advection_kind,
                   time_step_kind=E.time_step_kind, ts_params=E.ts_params,
                   dim_mask=E.dim_mask, backward=E.backward)


def release_ensemble(E, counts, cols, rows, lays, seed=1):
    """counts[m] particles of member m uniformly inside the block of cells,
    member-major (the order of the members' rng stream offsets)."""
    parts = []
    for m, n in enumerate(counts):
        c, x, y, z = release_in_block(E, int(n), cols, rows, lays, seed=seed + m)
        parts.append((E.global_cell(m, c).astype(np.int32), x, y, z))
    return tuple(np.concatenate([p[k] for p in parts]) for k in range(4))


def column_zones(P: Problem, nzones: int) -> np.ndarray:
    """0-based zone of every cell: nzones bands of equal column count along x."""
    j = np.arange(P.ncell) % P.ncol
    return (j * nzones // P.ncol).astype(np.int32)


def zoned_dispersion(P: Problem, zones: np.ndarray, aL_z, aT_z, dm) -> Problem:
    """P with per-cell aL, aT from zone values and a uniform Dm (per cell)."""
    Q = dataclasses.replace(P)
    Q.aL = np.asarray(aL_z, dtype=np.float64)[zones]
    Q.aT = np.asarray(aT_z, dtype=np.float64)[zones]
    Q.dmeff = np.full(P.ncell, float(dm))
    return Q


def config4(nzones=8, seed=20240601, **kw):
    """The calibration workload: the config-1 field with 8 x-bands of 25
    columns, each with its own aL and aT, a common Dm; returns the problem,
    the zone of every cell and the 18 parameters (aL_z, aT_z, porosity
    scale 1, Dm)."""
    P = config1(seed=seed, **kw)
    zones = column_zones(P, nzones)
    aL = np.linspace(0.3, 0.7, nzones)
    aT = 0.1 * aL
    return zoned_dispersion(P, zones, aL, aT, 1e-4), zones, (aL, aT, 1.0, 1e-4)


def observation_points(P: Problem, n_x=8, n_y=4, cols=(20, 160), rows=(85, 115)):
    """Global (x, y, z) of n_x * n_y points at cell centres of the middle
    layer along the plume path (x along columns, y along rows)."""
    xc = np.cumsum(P.delr) - P.delr / 2
    yc = np.cumsum(P.delc[::-1])[::-1] - P.delc / 2  # row 1 at the top (largest y)
    k = P.nlay // 2
    c = k * P.nrow * P.ncol
    zmid = 0.5 * (P.top[c] + P.bot[c])
    js = np.linspace(cols[0], cols[1], n_x).astype(int) - 1
    is_ = np.linspace(rows[0], rows[1], n_y).astype(int) - 1
    return np.array([(xc[j], yc[i], zmid) for i in is_ for j in js])


@dataclasses.dataclass
class UsgProblem:
    """An unstructured grid of rectangular cells (MODFLOW-USG, or a MODFLOW
    6 DISV smoothed quadtree) with its flows, as ModflowRectangularGrid and
    FlowModelData hold them (RectangularGridDisvMf6.f90:47-274,
    FlowModelData.f90:217-622):

    * ``ia`` (ncell+1) and ``ja`` are MODFLOW's 1-based CSR connectivity,
      each row led by the cell itself; ``layer`` is each cell's 1-based
      layer; ``cell_x, cell_y, del_x, del_y`` its centre and sizes;
    * ``flowja`` is FLOW-JA-FACE: entry q of row n is the flow into n from
      cell ja[q]; ``bflow`` (6 per cell), ``storage``, ``source``, ``sink``
      the boundary, storage and source/sink flows;
    * the per-cell properties and the options are those of ``Problem``;
      ``layer_type`` is the grid's ICELLTYPE (GetDZRW).
    """
    ncell: int
    nlay: int
    ia: np.ndarray
    ja: np.ndarray
    layer: np.ndarray
    cell_x: np.ndarray
    cell_y: np.ndarray
    del_x: np.ndarray
    del_y: np.ndarray
    top: np.ndarray
    bot: np.ndarray
    flowja: np.ndarray
    porosity: np.ndarray
    bflow: Optional[np.ndarray] = None
    storage: Optional[np.ndarray] = None
    source: Optional[np.ndarray] = None
    sink: Optional[np.ndarray] = None
    retardation: Optional[np.ndarray] = None
    head: Optional[np.ndarray] = None
    layer_type: Optional[np.ndarray] = None
    zone: Optional[np.ndarray] = None
    ibound: Optional[np.ndarray] = None
    icbound: Optional[np.ndarray] = None
    default_icbound: int = 0
    disp_model: int = 0
    aL: np.ndarray = dataclasses.field(default_factory=lambda: np.array([0.1]))
    aT: np.ndarray = dataclasses.field(default_factory=lambda: np.array([0.01]))
    aLV: Optional[np.ndarray] = None
    aTV: Optional[np.ndarray] = None
    dmeff: np.ndarray = dataclasses.field(default_factory=lambda: np.array([1e-4]))
    advection_kind: int = 2
    time_step_kind: int = 4
    ts_params: tuple = (0.25, 0.0)
    dim_mask: tuple = (1, 1, 1)
    backward: int = 0
    stop_zone: int = 0
    stop_weak_sinks: int = 0
    stop_weak_sources: int = 0
    steady_state: int = 1
    saturated_top: Optional[np.ndarray] = None   # as Problem.saturated_top

    @property
    def nja(self) -> int:
        return int(np.asarray(self.ja).size)

    @property
    def disp_uniform(self) -> int:
        return int(np.asarray(self.aL).size == 1)

    @property
    def uniform_porosity(self) -> int:
        """isUniformPorosity, ModpathBasicData.f90:208."""
        p = np.asarray(self.porosity)
        return int(bool(np.all(p == p[0])))
