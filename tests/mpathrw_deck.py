"""A MODPATH-RW run over a synthetic MODFLOW 6 DIS model, described once and
written twice: as the input deck MODPATH-RW reads (name file, basic data,
simulation, RW options, dispersion and gpkde files over the binary grid,
head and budget files of tests/mf6_synth.py and a TDIS file) and as the
arguments of rwpt.driver for the same run.  Test infrastructure.

The DIS model: a 14 x 10 x 3 heterogeneous divergence-free flow field
(problems.heterogeneous_flow) entering through the west face of column 1
and leaving through the east face of column 14 (CHD with an IFACE
auxiliary), a pumping well (a weak sink), a convertible top layer with its
water table inside the layer, three MODFLOW time steps in two stress
periods whose flow rates differ.  The DISV model (disv_run): a smoothed
quadtree of 12 x 8 base cells with a refined block and a refined pair in 3
layers (tests/usg_synth.py), its stream-function flow entering through the
east face and leaving through the west face, the same well, layer types
and time steps.  Dispersion is zero, so every path is fixed by the flow and
neither run's random draws matter.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Sequence

import numpy as np

from rwpt import problems

import mf6_synth as S


@dataclasses.dataclass
class Group:
    """A particle group of the simulation file (ModpathSimulationData.f90:
    635-720) with starting locations of input style 1
    (StartingLocationReader.f90:148-240)."""
    name: str
    release: Sequence[float]          # release option 1 (one time) or 3
    cells: np.ndarray                 # 1-based cell numbers
    lx: np.ndarray
    ly: np.ndarray
    lz: np.ndarray
    offset: np.ndarray                # release time offsets of the first release
    by_layer_row_column: bool = False  # location style 1 instead of 2
    ids: Optional[np.ndarray] = None  # particle IDs read from the file


@dataclasses.dataclass
class Run:
    nlay: int = 3
    nrow: int = 10
    ncol: int = 14
    seed: int = 11
    porosity: float = 0.25
    perlen: Sequence[float] = (6.0, 8.0)
    nstp: Sequence[int] = (1, 2)
    rate: Sequence[float] = (1.0, 1.5, 0.8)      # flow scale of each time step
    time_points: Sequence[float] = (2.5, 6.0, 9.0, 13.0, 17.5)
    stop_time: float = 20.0
    dt: float = 0.25                             # FIXED RW time step
    timeseries_option: int = 0                   # TimeseriesOutputOption
    gpkde_points: Optional[Sequence[float]] = (5.0, 10.0, 15.0)  # None: follow
    gpkde_skip_initial: bool = False
    groups: List[Group] = dataclasses.field(default_factory=list)
    grid: str = "dis"                            # or "disv"
    refine: tuple = ()                           # refined base cells (DISV)

    def problem(self):
        P = problems.heterogeneous_flow(self.ncol, self.nrow, self.nlay, corr_len=3.0,
                                        seed=self.seed)
        P.porosity = np.full(P.ncell, self.porosity)
        return P

    @property
    def step_ends(self):
        ends, t = [], 0.0
        for L, n in zip(self.perlen, self.nstp):
            for k in range(n):
                t += L / n
                ends.append(t)
        return ends

    @property
    def steps(self):
        return [(p + 1, s + 1) for p, n in enumerate(self.nstp) for s in range(n)]


def quadtree(R: Run):
    from usg_synth import Quadtree
    return Quadtree(R.nrow, R.ncol, set(R.refine), nlay=R.nlay)


def disv_run(**kw) -> Run:
    """The DISV run: particles enter with the flow near the east edge."""
    kw.setdefault("nrow", 8)
    kw.setdefault("ncol", 12)
    kw.setdefault("refine", ((2, 4), (2, 5), (3, 4), (3, 5), (5, 8), (6, 8)))
    R = Run(grid="disv", **kw)
    Q = quadtree(R)
    g = np.random.default_rng(R.seed + 1)
    east = np.nonzero(Q.x0 >= (R.ncol - 3) * 1.0)[0]
    layer_of = lambda p, k: k * Q.ncpl + p + 1
    n1 = 30
    cells = np.array([layer_of(int(p), int(k)) for p, k in
                      zip(g.choice(east, n1), g.integers(0, R.nlay, n1))], np.int32)
    off = np.zeros(n1)
    off[-6:] = 2.0
    R.groups.append(Group("east", [0.0], cells, np.round(g.uniform(0.05, 0.95, n1), 3),
                          np.round(g.uniform(0.05, 0.95, n1), 3),
                          np.round(g.uniform(0.05, 0.95, n1), 3), off))
    mid = np.nonzero((Q.x0 >= 4.0) & (Q.x1 <= 9.0))[0]
    n2 = 10
    cells = np.array([layer_of(int(p), int(k)) for p, k in
                      zip(g.choice(mid, n2), g.integers(0, R.nlay, n2))], np.int32)
    lx = np.round(g.uniform(0.05, 0.95, n2), 3)
    lx[0] = 1.0                         # on face 2
    R.groups.append(Group("twice", [3.0, 9.0], cells, lx, np.full(n2, 0.5),
                          np.full(n2, 0.25), np.zeros(n2),
                          ids=np.arange(201, 201 + n2, dtype=np.int32)))
    R.groups.append(Group("late", [30.0], cells[:2].copy(), np.full(2, 0.5),
                          np.full(2, 0.5), np.full(2, 0.5), np.zeros(2)))
    return R


def default_run(**kw) -> Run:
    R = Run(**kw)
    if R.groups:
        return R
    P = R.problem()
    g = np.random.default_rng(R.seed + 1)
    n1 = 40
    k = g.integers(1, R.nlay + 1, n1)
    i = g.integers(1, R.nrow + 1, n1)
    j = g.integers(1, 4, n1)
    off = np.zeros(n1)
    off[-8:] = 3.0                      # released inside the first step
    R.groups.append(Group("inlet", [0.0], P.cell_number(k, i, j).astype(np.int32),
                          np.round(g.uniform(0.05, 0.95, n1), 3),
                          np.round(g.uniform(0.05, 0.95, n1), 3),
                          np.round(g.uniform(0.05, 0.95, n1), 3), off))
    # a second group released twice, read by layer, row and column with IDs
    n2 = 12
    k = g.integers(1, R.nlay + 1, n2)
    i = g.integers(1, R.nrow + 1, n2)
    j = g.integers(2, 7, n2)
    lx = np.round(g.uniform(0.05, 0.95, n2), 3)
    lx[0] = 0.0                         # on face 1
    R.groups.append(Group("twice", [4.0, 11.0], P.cell_number(k, i, j).astype(np.int32),
                          lx, np.full(n2, 0.5), np.full(n2, 1.0), np.zeros(n2),
                          by_layer_row_column=True,
                          ids=np.arange(101, 101 + n2, dtype=np.int32)))
    # never released within the run
    R.groups.append(Group("late", [30.0], P.cell_number(np.full(3, 2), np.array([3, 5, 7]),
                                                        np.full(3, 2)).astype(np.int32),
                          np.full(3, 0.5), np.full(3, 0.5), np.full(3, 0.5), np.zeros(3)))
    return R


WELL = (2, 5, 9)       # layer, row, column of the pumping well


def flows(R: Run, P):
    """FLOW-JA-FACE and budget records of each time step, and the heads."""
    ncpl = R.nrow * R.ncol
    n = P.ncell
    icelltype = np.zeros(n, np.int32)
    icelltype[:ncpl] = 1
    g = np.random.default_rng(R.seed + 2)
    head = P.top.copy() + 1.0
    head[:ncpl] = P.bot[:ncpl] + (P.top - P.bot)[:ncpl] * g.uniform(0.55, 0.95, ncpl)
    west = np.array([int(P.cell_number(k, i, 1)) for k in range(1, R.nlay + 1)
                     for i in range(1, R.nrow + 1)])
    east = west + R.ncol - 1
    well = int(P.cell_number(*WELL))
    out = []
    for s in R.rate:
        qin = float(P.bflow[6 * (west[0] - 1)]) * s
        recs = [("list", "CHD", "CHD-1", list(west) + list(east),
                 [qin] * len(west) + [-qin] * len(east), ("IFACE",),
                 [[1.0]] * len(west) + [[2.0]] * len(east)),
                ("list", "WEL", "WEL-1", [well], [-0.05 * s], (), None)]
        out.append((s, recs))
    return icelltype, head, out


def disv_flows(R: Run, Q):
    """flows() for the DISV model: the quadtree's FLOW-JA-FACE scaled per
    step, CHD with IFACE on the west and east edges, the well."""
    n, ncpl = Q.ncell, Q.ncpl
    icelltype = np.zeros(n, np.int32)
    icelltype[:ncpl] = 1
    g = np.random.default_rng(R.seed + 2)
    top = Q.top
    head = top + 1.0
    head[:ncpl] = Q.botm[:ncpl] + (top - Q.botm)[:ncpl] * g.uniform(0.55, 0.95, ncpl)
    fja = Q.flows(seed=R.seed)
    bnd = Q.boundary_flows()
    well = (WELL[0] - 1) * ncpl + int(np.nonzero((Q.x0 <= 5.5) & (Q.x1 > 5.5) &
                                                 (Q.y0 <= 4.5) & (Q.y1 > 4.5))[0][0]) + 1
    out = []
    for s in R.rate:
        recs = [("list", "CHD", "CHD-1", [c for c, _, _ in bnd], [q * s for _, _, q in bnd],
                 ("IFACE",), [[float(f)] for _, f, _ in bnd]),
                ("list", "WEL", "WEL-1", [well], [-0.05 * s], (), None)]
        out.append((s, fja * s, recs))
    return icelltype, head, out


def write(R: Run, tmp: str, name: str = "m") -> dict:
    """Writes the MODFLOW 6 files and MODPATH-RW's deck into tmp; returns
    the paths (the simulation file first)."""
    disv = R.grid == "disv"
    p = {k: os.path.join(tmp, f"{name}.{k}") for k in
         ("grb", "hds", "bud", "tdis", "mpnam", "mpbas", "mpsim", "rwopts", "dsp", "gpkde",
          "imp")}
    out = {k: os.path.join(tmp, f"{name}.{k}") for k in ("mpend", "timeseries", "gpkde.out",
                                                          "mplst")}
    totim = R.step_ends
    nlay = R.nlay
    if disv:
        Q = quadtree(R)
        ncpl = Q.ncpl
        icelltype, head, steps = disv_flows(R, Q)
        Q.write_grb(p["grb"], icelltype=icelltype)
        hrow, hcol = 1, ncpl
        bot_min, xsize, ysize = float(Q.botm.min()), R.ncol * 1.0, R.nrow * 1.0
        zsize = float(Q.top.max() - Q.botm.min())
    else:
        P = R.problem()
        nrow, ncol = R.nrow, R.ncol
        ncpl = nrow * ncol
        icelltype, head, recs_of = flows(R, P)
        ia, ja = S.write_grb(p["grb"], nlay, nrow, ncol, P.delr, P.delc, P.top[:ncpl], P.bot,
                             icelltype)
        steps = [(s, S.flow_ja(ia, ja, P.frf * s, P.fff * s, P.flf * s, nrow, ncol), recs)
                 for s, recs in recs_of]
        hrow, hcol = nrow, ncol
        bot_min, xsize, ysize = float(P.bot.min()), float(P.delr.sum()), float(P.delc.sum())
        zsize = float(P.top.max() - P.bot.min())
    S.write_heads(p["hds"], nlay, hrow, hcol, head,
                  steps=[(kp, ks, t) for (kp, ks), t in zip(R.steps, totim)])
    with open(p["bud"], "wb") as f:
        for (kp, ks), t, (scale, fja, recs) in zip(R.steps, totim, steps):
            tmpf = p["bud"] + ".part"
            S.write_budget(tmpf, nlay, hrow, hcol, fja, recs, kper=kp, kstp=ks, totim=t)
            with open(tmpf, "rb") as g:
                f.write(g.read())
            os.remove(tmpf)
    with open(p["tdis"], "w") as f:
        f.write("BEGIN options\n  TIME_UNITS days\nEND options\n\n")
        f.write(f"BEGIN dimensions\n  NPER {len(R.perlen)}\nEND dimensions\n\n")
        f.write("BEGIN perioddata\n")
        for L, n in zip(R.perlen, R.nstp):
            f.write(f"  {L:.8f} {n} 1.0\n")
        f.write("END perioddata\n")
    b = os.path.basename
    with open(p["mpnam"], "w") as f:
        f.write("# MODPATH-RW name file\n")
        for key, k in (("MPBAS", "mpbas"), ("GRBDISV" if disv else "GRBDIS", "grb"),
                       ("TDIS", "tdis"),
                       ("HEAD", "hds"), ("BUDGET", "bud"), ("DSP", "dsp"),
                       ("RWOPTS", "rwopts"), ("GPKDE", "gpkde"), ("IMP", "imp")):
            f.write(f"{key:<11}{b(p[k])}\n")
    with open(p["mpbas"], "w") as f:
        f.write("# basic data\n0\n")
        for _ in range(nlay):
            f.write(f"CONSTANT {R.porosity}\n")
    with open(p["rwopts"], "w") as f:
        f.write(f"FIXED\n{R.dt}\nEULERIAN\n1 1 1\n")
    # open model boundaries (defaultICBound 0) and ICBOUND from IBOUND
    # (TransportModelData.f90:209-240); without this file the reference
    # reflects particles at every model boundary
    with open(p["imp"], "w") as f:
        f.write("0\n0\n")
    with open(p["dsp"], "w") as f:
        f.write("1\nDSP1\n0 0\n0.0\n0.0\n0.0\n")
    with open(p["gpkde"], "w") as f:
        f.write(f"{b(out['gpkde.out'])} 0 0\n")
        f.write(f"0.0 0.0 {bot_min}\n")
        f.write(f"{xsize} {ysize} {zsize}\n")
        f.write("1.0 1.0 1.0\n0\n0\n0 0.02\n0 0 0 2\n0\n0\n0\n")
        skip = 1 if R.gpkde_skip_initial else 0
        if R.gpkde_points is None:
            f.write(f"0 {skip}\n")
        else:
            f.write(f"2 {skip}\n{len(R.gpkde_points)}\n")
            f.write(" ".join(repr(float(t)) for t in R.gpkde_points) + "\n")
    with open(p["mpsim"], "w") as f:
        f.write("# MODPATH-RW simulation file\n")
        f.write(f"{b(p['mpnam'])}\n{b(out['mplst'])}\n")
        # type 5, forward, stop at weak sinks, pass weak sources, no budget
        # output, no trace, the timeseries option, endpoints, no masses,
        # one solute
        f.write(f"5 1 2 1 0 0 {R.timeseries_option} 0 0 0\n")
        f.write(f"{b(out['mpend'])}\n{b(out['timeseries'])}\n")
        f.write("0\n1\n0.0\n3\n" + repr(float(R.stop_time)) + "\n")
        f.write(f"2\n{len(R.time_points)}\n")
        f.write(" ".join(repr(float(t)) for t in R.time_points) + "\n")
        f.write("1\n1\n")
        f.write(f"{len(R.groups)}\n")
        for G in R.groups:
            f.write(f"{G.name}\n")
            if len(G.release) == 1:
                f.write(f"1\n{float(G.release[0])!r}\n")
            else:
                f.write(f"3\n{len(G.release)}\n")
                f.write(" ".join(repr(float(t)) for t in G.release) + "\n")
            f.write("INTERNAL\n1\n")
            f.write("1\n" if G.by_layer_row_column and not disv else "2\n")
            n = len(G.cells)
            f.write(f"{n} {0 if G.ids is None else 1}\n")
            for m in range(n):
                c = int(G.cells[m])
                if G.by_layer_row_column and not disv:
                    k0, r = divmod(c - 1, ncpl)
                    loc = f"{k0 + 1} {r // ncol + 1} {r % ncol + 1}"
                else:
                    loc = f"{c}"
                pid = "" if G.ids is None else f"{int(G.ids[m])} "
                f.write(f"{pid}{loc} {float(G.lx[m])!r} {float(G.ly[m])!r} "
                        f"{float(G.lz[m])!r} {float(G.offset[m])!r} 0\n")
    return dict(sim=p["mpsim"], files=p, out=out)
