"""MODFLOW 6 DIS output as engine input.

MPathRW reads three files of a MODFLOW 6 simulation with a structured (DIS)
grid: the binary grid file, the head file and the cell-by-cell budget file
(MPathRW.F90:305-330, 394-433).  This module reads them and turns one time
step into the arrays of rwpt_set_flow and rwpt_set_cell_extras, as the
reference's MODFLOW 6 path does:

- ``read_grb`` follows ReadGRB (RectangularGridDisMf6.f90:122-330): four
  50-character header lines (GRID DIS, VERSION 1, NTXT 16, LENTXT 100), 16
  definition lines of 100 characters whose first word names the variable,
  then the variables in that order; the top of every cell below the first
  layer is the bottom of the cell above.
- ``read_heads`` reads the structured head records HeadReader takes from a
  MODFLOW 6 DIS head file (kstp, kper, pertim, totim, text, ncol, nrow,
  ilay and one layer of doubles).
- ``read_budget`` reads compact budget records as BudgetReader's record
  header does (BudgetReader.f90:560-720): kstp, kper, text, ndim1..3, and,
  with ndim3 < 0, method, delt, pertim, totim; method 1 arrays, method 2
  lists, method 5 and 6 lists with auxiliary values (method 6 with the four
  16-character model and package names and two ids per item).
- ``load_time_step`` restates FlowModelData%LoadTimeStep for MODFLOW 6 grids
  (FlowModelData.f90:217-622): IBoundTS from the dry-cell test of
  ICELLTYPE 1 cells (single-precision comparison with HDry), then each
  budget term of the time step into FlowsJA, StorageFlows, BoundaryFlows
  (IFACE from the IFACE auxiliary variable or the default IFACE labels)
  or the source and sink flows.
- ``engine_arrays`` gives the engine's MODFLOW-2005 face flows: on a DIS
  grid a face has at most one connection, and SetMassTransportDataUnstructured
  (ModpathCellData.f90:886-950) sets Q1 = F(n, left), Q2 = -F(n, right),
  Q3 = F(n, front), Q4 = -F(n, back), Q5 = F(n, below), Q6 = -F(n, above)
  from FLOW-JA-FACE F.  The structured assembly the engine implements
  (ModpathCellData.f90:1315-1351) forms Q2 = frf(n), Q3 = -fff(n),
  Q5 = -flf(n) and Q1, Q4, Q6 from the neighbours' values, so
  frf(n) = -F(n, right), fff(n) = -F(n, front), flf(n) = -F(n, below);
  MODFLOW 6 writes F antisymmetric (F(m, n) = -F(n, m)), which makes the
  neighbours' faces agree as well.

Rows grow toward the front (the y face coordinates fall with the row,
RectangularGridDisMf6.f90:300-306) and layers downward.

Cells with IDOMAIN < 1 are removed from the model: the binary grid file's
IA and JA use the user cell numbers with an empty row for each removed cell
and no connection to it, FLOW-JA-FACE follows JA position by position, list
budget records name user cells and array records cover every user cell.
MODPATH sets IBOUND = 0 at such cells (ModpathBasicData.f90:177-185,
ibound_from_idomain); the engine's face flow across a missing connection
is 0, as the unstructured assembly leaves a face without a connection.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_I4 = np.dtype("<i4")
_F8 = np.dtype("<f8")


class Mf6Error(ValueError):
    """A file that is not MODFLOW 6 DIS output of the supported form."""


class _Stream:
    """Sequential reads of a binary file, with skips that do not read."""

    def __init__(self, path):
        import os
        self.f = open(path, "rb")
        self.size = os.fstat(self.f.fileno()).st_size

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def eof(self):
        return self.f.tell() >= self.size

    def take(self, n):
        b = self.f.read(n)
        if len(b) != n:
            raise Mf6Error("unexpected end of file")
        return b

    def skip(self, n):
        if self.f.tell() + n > self.size:
            raise Mf6Error("unexpected end of file")
        self.f.seek(n, 1)

    def i4(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.take(4 * k), _I4)
        return int(a[0]) if n is None else a.astype(np.int32)

    def f8(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.take(8 * k), _F8)
        return float(a[0]) if n is None else a.astype(np.float64)

    def text(self, n):
        return self.take(n).decode("ascii", "replace")


@dataclasses.dataclass
class Grb:
    ncells: int
    nlay: int
    nrow: int
    ncol: int
    nja: int
    xorigin: float
    yorigin: float
    angrot: float
    delr: np.ndarray
    delc: np.ndarray
    top: np.ndarray        # per cell (the layer above's bottom below layer 1)
    botm: np.ndarray
    ia: np.ndarray         # 0-based offsets, ncells + 1
    ja: np.ndarray         # 1-based cell numbers
    idomain: np.ndarray
    icelltype: np.ndarray


@dataclasses.dataclass
class GrbDisv:
    """A DISV binary grid of rectangular cells as RectangularGridDisvMf6
    holds it after ReadGRB and CheckRectangular: per cell (all layers) the
    centre and sizes from the vertices, tops, bottoms, IA (0-based offsets)
    and JA, IDOMAIN and ICELLTYPE."""
    ncells: int
    nlay: int
    ncpl: int
    nja: int
    xorigin: float
    yorigin: float
    angrot: float
    top: np.ndarray
    botm: np.ndarray
    ia: np.ndarray
    ja: np.ndarray
    idomain: np.ndarray
    icelltype: np.ndarray
    cell_x: np.ndarray
    cell_y: np.ndarray
    del_x: np.ndarray
    del_y: np.ndarray

    @property
    def layer(self) -> np.ndarray:
        """1-based layer of every cell (GetLayer over LayerOffsets)."""
        return (np.arange(self.ncells) // self.ncpl + 1).astype(np.int32)


def _ncpl(grb) -> int:
    return grb.ncpl if isinstance(grb, GrbDisv) else grb.nrow * grb.ncol


def read_grb(path):
    """The binary grid file of a DIS model (ReadGRB,
    RectangularGridDisMf6.f90:122-330) as a Grb, or of a DISV model of
    rectangular cells (RectangularGridDisvMf6.f90:47-274) as a GrbDisv."""
    with _Stream(path) as s:
        words = s.text(50).split()
        if words[:2] == ["GRID", "DISV"]:
            return _read_grb_disv(s)
        if words[:2] != ["GRID", "DIS"]:
            raise Mf6Error("not a DIS or DISV binary grid file")
        return _read_grb(s)


def _colinear(x1, y1, x2, y2, x3, y3, tol):
    # IsColinear (RectangularGridDisvMf6.f90:309-323)
    d1 = np.sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1))
    d2 = np.sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2))
    d3 = np.sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1))
    return abs((d3 - d1 - d2) / d3) <= tol


# the four clockwise side directions of a rectangle starting at each corner
# (CheckRectangular's ix, iy tables)
_SIDES = [((1, 0), (0, -1), (-1, 0), (0, 1)), ((0, -1), (-1, 0), (0, 1), (1, 0)),
          ((-1, 0), (0, 1), (1, 0), (0, -1)), ((0, 1), (1, 0), (0, -1), (-1, 0))]


def _check_rectangular(ncpl, iavert, javert, vx, vy):
    """CheckRectangular (RectangularGridDisvMf6.f90:326-570): each cell's
    corners (vertices where the outline turns), at most one mid-face vertex
    per side, four axis-aligned sides; the corners reordered from the
    upper-left one; DelX, DelY, CellX, CellY of layer 1 from them."""
    tol = float(np.float32(0.00001))
    smoothed = True
    dx, dy, cx, cy = (np.zeros(ncpl) for _ in range(4))
    for n in range(ncpl):
        off, cnt = iavert[n], iavert[n + 1] - iavert[n]
        corners = [javert[off]]
        pts, mid = 1, 0
        for i in range(1, cnt - 1):
            v1, v2, v3 = javert[off + i - 1], javert[off + i], javert[off + i + 1]
            if not _colinear(vx[v1 - 1], vy[v1 - 1], vx[v2 - 1], vy[v2 - 1], vx[v3 - 1],
                             vy[v3 - 1], tol):
                pts += 1
                if pts <= 4:
                    corners.append(v2)
                if mid > 1:
                    smoothed = False
                mid = 0
            else:
                mid += 1
        pts += 1
        if pts != 5:
            raise Mf6Error("a DISV cell is not a quadrilateral")
        dirs = []
        for a in range(4):
            p, q = corners[a] - 1, corners[(a + 1) % 4] - 1
            ddx, ddy = vx[q] - vx[p], vy[q] - vy[p]
            dirs.append((1 if ddx > tol else (-1 if ddx < -tol else 0),
                         1 if ddy > tol else (-1 if ddy < -tol else 0)))
        k = next((m for m in range(4) if tuple(dirs) == _SIDES[m]), -1)
        if k < 0:
            raise Mf6Error("a DISV cell is not a rectangle aligned with the axes")
        corners = corners[k:] + corners[:k]
        x1, x2 = vx[corners[0] - 1], vx[corners[1] - 1]
        y1, y2 = vy[corners[3] - 1], vy[corners[0] - 1]
        dx[n] = x2 - x1
        cx[n] = (x1 + x2) / 2.0
        dy[n] = y2 - y1
        cy[n] = (y1 + y2) / 2.0
    if not smoothed:
        raise Mf6Error("a DISV cell side meets more than two cells")
    return dx, dy, cx, cy


def _read_grb_disv(s) -> GrbDisv:
    words = s.text(50).split()
    if words[:1] != ["VERSION"] or int(words[1]) != 1:
        raise Mf6Error("unsupported binary grid file version")
    words = s.text(50).split()
    if words[:1] != ["NTXT"]:
        raise Mf6Error("binary grid file lacks NTXT")
    ntxt = int(words[1])
    words = s.text(50).split()
    if words[:1] != ["LENTXT"] or int(words[1]) != 100:
        raise Mf6Error("binary grid file LENTXT must be 100")
    names = []
    for _ in range(ntxt):
        line = s.text(100)
        names.append("#" if line.startswith("#") else line.split()[0])
    v: Dict[str, object] = {}
    for name in names:
        if name == "#":
            continue
        if name in ("NCELLS", "NLAY", "NCPL", "NVERT", "NJAVERT", "NJA"):
            v[name] = s.i4()
        elif name in ("XORIGIN", "YORIGIN", "ANGROT"):
            v[name] = s.f8()
        elif name in ("TOP", "CELLX", "CELLY"):
            v[name] = s.f8(v["NCPL"])
        elif name == "BOTM":
            v[name] = s.f8(v["NCELLS"])
        elif name == "VERTICES":
            v[name] = s.f8(2 * v["NVERT"]).reshape(-1, 2)
        elif name == "IAVERT":
            v[name] = s.i4(v["NCPL"] + 1) - 1
        elif name == "JAVERT":
            v[name] = s.i4(v["NJAVERT"])
        elif name == "IA":
            v[name] = s.i4(v["NCELLS"] + 1) - 1
        elif name == "JA":
            v[name] = s.i4(v["NJA"])
        elif name in ("IDOMAIN", "ICELLTYPE"):
            v[name] = s.i4(v["NCELLS"])
        else:
            raise Mf6Error(f"unknown binary grid variable {name}")
    need = ("NCELLS", "NLAY", "NCPL", "NVERT", "NJAVERT", "NJA", "TOP", "BOTM",
            "VERTICES", "IAVERT", "JAVERT", "IA", "JA", "IDOMAIN", "ICELLTYPE")
    if any(k not in v for k in need):
        raise Mf6Error("binary grid file lacks a required variable")
    n, ncpl, nlay = v["NCELLS"], v["NCPL"], v["NLAY"]
    if n != nlay * ncpl:
        raise Mf6Error("NCELLS differs from NLAY * NCPL")
    verts = v["VERTICES"]
    dx, dy, cx, cy = _check_rectangular(ncpl, v["IAVERT"], v["JAVERT"], verts[:, 0],
                                        verts[:, 1])
    top = np.empty(n)
    top[:ncpl] = v["TOP"]
    top[ncpl:] = v["BOTM"][:n - ncpl]
    rep = lambda a: np.tile(a, nlay)
    return GrbDisv(n, nlay, ncpl, v["NJA"], v.get("XORIGIN", 0.0), v.get("YORIGIN", 0.0),
                   v.get("ANGROT", 0.0), top, v["BOTM"], v["IA"], v["JA"], v["IDOMAIN"],
                   v["ICELLTYPE"], rep(cx), rep(cy), rep(dx), rep(dy))


def _read_grb(s) -> Grb:
    words = s.text(50).split()
    if words[:1] != ["VERSION"] or int(words[1]) != 1:
        raise Mf6Error("unsupported binary grid file version")
    words = s.text(50).split()
    if words[:1] != ["NTXT"] or int(words[1]) != 16:
        raise Mf6Error("binary grid file NTXT must be 16")
    words = s.text(50).split()
    if words[:1] != ["LENTXT"] or int(words[1]) != 100:
        raise Mf6Error("binary grid file LENTXT must be 100")
    names = []
    for _ in range(16):
        line = s.text(100)
        names.append("#" if line.startswith("#") else line.split()[0])
    v: Dict[str, object] = {}
    for name in names:
        if name == "#":
            continue
        if name in ("NCELLS", "NLAY", "NROW", "NCOL", "NJA"):
            v[name] = s.i4()
        elif name in ("XORIGIN", "YORIGIN", "ANGROT"):
            v[name] = s.f8()
        elif name == "DELR":
            v[name] = s.f8(v["NCOL"])
        elif name == "DELC":
            v[name] = s.f8(v["NROW"])
        elif name == "TOP":
            v[name] = s.f8(v["NROW"] * v["NCOL"])
        elif name == "BOTM":
            v[name] = s.f8(v["NCELLS"])
        elif name == "IA":
            v[name] = s.i4(v["NCELLS"] + 1) - 1
        elif name == "JA":
            v[name] = s.i4(v["NJA"])
        elif name in ("IDOMAIN", "ICELLTYPE"):
            v[name] = s.i4(v["NCELLS"])
        else:
            raise Mf6Error(f"unknown binary grid variable {name}")
    need = ("NCELLS", "NLAY", "NROW", "NCOL", "NJA", "DELR", "DELC", "TOP", "BOTM", "IA",
            "JA", "IDOMAIN", "ICELLTYPE")
    if any(k not in v for k in need):
        raise Mf6Error("binary grid file lacks a required variable")
    n, ncpl = v["NCELLS"], v["NROW"] * v["NCOL"]
    if n != v["NLAY"] * ncpl:
        raise Mf6Error("NCELLS differs from NLAY * NROW * NCOL")
    top = np.empty(n)
    top[:ncpl] = v["TOP"]
    top[ncpl:] = v["BOTM"][:n - ncpl]
    return Grb(n, v["NLAY"], v["NROW"], v["NCOL"], v["NJA"], v.get("XORIGIN", 0.0),
               v.get("YORIGIN", 0.0), v.get("ANGROT", 0.0), v["DELR"], v["DELC"], top,
               v["BOTM"], v["IA"], v["JA"], v["IDOMAIN"], v["ICELLTYPE"])


@dataclasses.dataclass
class HeadRecord:
    kstp: int
    kper: int
    pertim: float
    totim: float
    text: str
    ncol: int
    nrow: int
    ilay: int
    data: np.ndarray


def read_heads(path, kper=None, kstp=None) -> List[HeadRecord]:
    """Structured double-precision head records (of one time step when kper
    and kstp are given; the others are skipped unread)."""
    out = []
    with _Stream(path) as s:
        while not s.eof():
            kstp_, kper_ = s.i4(), s.i4()
            pertim, totim = s.f8(), s.f8()
            text = s.text(16)
            ncol, nrow, ilay = s.i4(), s.i4(), s.i4()
            if ncol < 1 or nrow < 1 or ilay < 1:
                raise Mf6Error("not a structured head file")
            if (kper is not None and kper_ != kper) or (kstp is not None and kstp_ != kstp):
                s.skip(8 * ncol * nrow)
                continue
            out.append(HeadRecord(kstp_, kper_, pertim, totim, text, ncol, nrow, ilay,
                                  s.f8(ncol * nrow)))
    return out


@dataclasses.dataclass
class BudgetRecord:
    kstp: int
    kper: int
    text: str                       # trimmed label
    ndim1: int
    ndim2: int
    ndim3: int
    method: int
    delt: float = 0.0
    pertim: float = 0.0
    totim: float = 0.0
    array: Optional[np.ndarray] = None     # methods 0, 1, 3, 4
    layer: Optional[np.ndarray] = None     # method 3
    node: Optional[np.ndarray] = None      # methods 2, 5, 6 (id1)
    node2: Optional[np.ndarray] = None     # method 6 (id2)
    q: Optional[np.ndarray] = None
    aux_names: Tuple[str, ...] = ()
    aux: Optional[np.ndarray] = None       # (nlist, naux)

    @property
    def array_item_count(self):
        if self.method in (2, 5, 6):
            return 0
        if self.method in (3, 4):
            return self.ndim1 * self.ndim2
        return self.ndim1 * self.ndim2 * self.ndim3


def read_budget(path, kper=None, kstp=None) -> List[BudgetRecord]:
    """Double-precision budget records, standard or compact (of one time
    step when kper and kstp are given; the others' data are skipped
    unread)."""
    with _Stream(path) as s:
        return _read_budget(s, kper, kstp)


def _read_budget(s, want_kper, want_kstp) -> List[BudgetRecord]:
    out = []
    while not s.eof():
        kstp, kper = s.i4(), s.i4()
        text = s.text(16).strip()
        n1, n2, n3 = s.i4(), s.i4(), s.i4()
        if n3 < 0:
            n3 = -n3
            method = s.i4()
            delt, pertim, totim = s.f8(), s.f8(), s.f8()
        else:
            method, delt, pertim, totim = 0, 0.0, 0.0, 0.0
        if not 0 <= method <= 6:
            raise Mf6Error(f"budget method {method}")
        r = BudgetRecord(kstp, kper, text, n1, n2, n3, method, delt, pertim, totim)
        keep = (want_kper is None or kper == want_kper) and \
            (want_kstp is None or kstp == want_kstp)
        if not keep and method in (0, 1, 3, 4):
            s.skip({0: 8 * n1 * n2 * n3, 1: 8 * n1 * n2 * n3, 3: 12 * n1 * n2,
                    4: 8 * n1 * n2}[method])
            continue
        if method in (0, 1):
            r.array = s.f8(n1 * n2 * n3)
        elif method == 3:
            r.layer = s.i4(n1 * n2)
            r.array = s.f8(n1 * n2)
        elif method == 4:
            r.array = s.f8(n1 * n2)
        else:
            naux = 0
            if method == 6:
                s.take(64)
            if method in (5, 6):
                nval = s.i4()
                naux = nval - 1
                r.aux_names = tuple(s.text(16).strip() for _ in range(naux))
            nlist = s.i4()
            ids = 2 if method == 6 else 1
            rec = np.dtype([("id", "<i4", (ids,)), ("v", "<f8", (1 + naux,))])
            if not keep:
                s.skip(rec.itemsize * nlist)
                continue
            a = np.frombuffer(s.take(rec.itemsize * nlist), rec)
            r.node = a["id"][:, 0].astype(np.int64)
            r.node2 = a["id"][:, 1].astype(np.int64) if ids == 2 else None
            r.q = a["v"][:, 0].astype(np.float64)
            r.aux = a["v"][:, 1:].astype(np.float64)
        out.append(r)
    return out


@dataclasses.dataclass
class TimeStepData:
    """FlowModelData's arrays after LoadTimeStep (1-based cells as index + 1)."""
    ibound_ts: np.ndarray
    heads: np.ndarray
    flows_ja: np.ndarray
    boundary_flows: np.ndarray      # 6 per cell, faces 1..6
    source_flows: np.ndarray
    sink_flows: np.ndarray
    storage_flows: np.ndarray
    steady_state: bool
    saturated_top: np.ndarray = None   # Grid%SaturatedTop


def _dry_tolerance(hdry):
    # HDryTol = abs(epsilon(HDryTol) * sngl(HDry)) in default REAL
    return np.float32(abs(np.finfo(np.float32).eps * np.float32(hdry)))


def load_time_step(grb, heads: Sequence[HeadRecord], budget: Sequence[BudgetRecord],
                   kper: int, kstp: int, ibound: np.ndarray, hdry: float,
                   default_iface: Dict[str, int] = None) -> TimeStepData:
    """FlowModelData%LoadTimeStep(kper, kstp) for a MODFLOW 6 DIS or DISV
    grid (a DISV head record holds one layer as NCPL columns of one row);
    ibound None takes MODPATH's IBOUND from IDOMAIN (ibound_from_idomain)."""
    n = grb.ncells
    default_iface = {k.strip(): int(v) for k, v in (default_iface or {}).items()}
    recs = [r for r in budget if r.kper == kper and r.kstp == kstp]
    if not recs:
        raise Mf6Error(f"no budget records for period {kper} step {kstp}")
    h = np.zeros(n)
    ncpl = _ncpl(grb)
    for r in heads:
        if r.kper == kper and r.kstp == kstp:
            if r.ncol * r.nrow != ncpl or not 1 <= r.ilay <= grb.nlay:
                raise Mf6Error("head record does not match the grid")
            h[(r.ilay - 1) * ncpl:r.ilay * ncpl] = r.data
    ib = np.asarray(ibound_from_idomain(grb) if ibound is None else ibound, np.int32).copy()
    # IBoundTS: dry ICELLTYPE 1 cells (FlowModelData.f90:260-282)
    conv = grb.icelltype == 1
    diff = np.abs(h.astype(np.float32) - np.float32(hdry))
    dry = conv & (diff < _dry_tolerance(hdry))
    ib[dry] = 0
    # SaturatedTop (FlowModelData.f90:259-280, MODFLOW 6 grids): the bottom
    # of a dry cell whose head lies below it, the head of a wet convertible
    # cell whose head lies between its bottom and top, else the top
    top = np.asarray(grb.top, np.float64)
    bot = np.asarray(grb.botm, np.float64)
    sat = top.copy()
    low = dry & (h < bot)
    sat[low] = bot[low]
    wet = conv & (ib != 0) & (h <= top) & (h >= bot)
    sat[wet] = h[wet]
    src, snk, sto = np.zeros(n), np.zeros(n), np.zeros(n)
    bfl = np.zeros(6 * n)
    fja = np.zeros(grb.nja)
    steady = True

    # np.add.at applies repeated indices one after another in item order,
    # so every sum is formed in the reference's order
    def split(cells0, values):
        c = np.asarray(cells0, np.int64)
        v = np.asarray(values, np.float64)
        pos, neg = v > 0.0, v < 0.0
        np.add.at(src, c[pos], v[pos])
        np.add.at(snk, c[neg], v[neg])

    def boundary(cells0, iface, values):
        c = np.asarray(cells0, np.int64)
        np.add.at(bfl, 6 * c + np.asarray(iface, np.int64) - 1,
                  np.asarray(values, np.float64))

    def listed(r):
        c = r.node - 1
        if "IFACE" in r.aux_names:
            # int() of the auxiliary value: truncation toward zero
            iface = np.trunc(r.aux[:, r.aux_names.index("IFACE")]).astype(np.int64)
        else:
            iface = np.full(len(c), default_iface.get(r.text, 0), np.int64)
        on = iface > 0
        # one cell's items go to a face or to source/sink in item order, and
        # the two targets are separate arrays, so the split keeps each order
        boundary(c[on], iface[on], r.q[on])
        split(c[~on], r.q[~on])

    for r in recs:
        t = r.text
        if t in ("CONSTANT HEAD", "CHD"):
            if r.method == 0:
                if r.array_item_count == n:
                    split(np.arange(n), r.array)
            elif r.method == 2:
                split(r.node - 1, r.q)
            elif r.method in (5, 6):
                listed(r)
        elif t in ("STORAGE", "STO-SS", "STO-SY"):
            if r.method in (0, 1) and r.array_item_count == n:
                sto += r.array
                if np.any(sto != 0.0):
                    steady = False
        elif t in ("FLOW JA FACE", "FLOW-JA-FACE"):
            if r.method in (0, 1):
                if r.array_item_count == grb.nja:
                    fja = r.array.copy()
            elif r.method == 6:
                split(r.node - 1, r.q)
        elif t in ("FLOW RIGHT FACE", "FLOW FRONT FACE", "FLOW LOWER FACE"):
            raise Mf6Error(f"{t} is a MODFLOW-2005 budget term")
        elif t in ("DATA-SPDIS", "DATA-SAT"):
            continue
        else:
            iface = default_iface.get(t, 0)
            if r.method in (0, 1):
                if r.array_item_count == n:
                    if iface > 0:
                        boundary(np.arange(n), iface, r.array)
                    else:
                        split(np.arange(n), r.array)
            elif r.method == 3:
                cells = (r.layer.astype(np.int64) - 1) * len(r.array) + \
                    np.arange(len(r.array))
                if iface > 0:
                    boundary(cells, iface, r.array)
                else:
                    split(r.layer.astype(np.int64) - 1, r.array)
            elif r.method == 4:
                if iface > 0:
                    boundary(np.arange(len(r.array)), iface, r.array)
                else:
                    split(np.arange(len(r.array)), r.array)
            elif r.method == 2:
                if iface > 0:
                    boundary(r.node - 1, iface, r.q)
                else:
                    split(r.node - 1, r.q)
            elif r.method in (5, 6):
                listed(r)
    return TimeStepData(ib, h, fja, bfl, src, snk, sto, steady, sat)


def ibound_from_idomain(grb: Grb) -> np.ndarray:
    """MODPATH's IBOUND of a MODFLOW 6 grid: 0 where IDOMAIN < 1, else 1
    (ModpathBasicData.f90:177-185)."""
    return np.where(grb.idomain < 1, 0, 1).astype(np.int32)


def _neighbour_index(grb: Grb):
    """Index into FLOW-JA-FACE of each cell's right, front and lower
    connection (-1 where the grid has no such neighbour or either cell is
    removed, IDOMAIN < 1)."""
    n, ncol, ncpl = grb.ncells, grb.ncol, grb.nrow * grb.ncol
    cnt = np.diff(grb.ia)
    live = grb.idomain >= 1
    if np.any(cnt[live] < 1) or np.any(cnt[~live] != 0) or \
            np.any(grb.ja[grb.ia[:-1][live]] - 1 != np.flatnonzero(live)):
        raise Mf6Error("JA must list each cell first in its own row "
                       "(and nothing for a removed cell)")
    # (row cell, connected cell) pairs as one sortable key
    row = np.repeat(np.arange(n, dtype=np.int64), cnt)
    key = row * (n + 1) + (grb.ja.astype(np.int64) - 1)
    order = np.argsort(key, kind="stable")
    skey = key[order]
    cell = np.arange(n, dtype=np.int64)
    col, rrow, lay = cell % ncol, (cell % ncpl) // ncol, cell // ncpl
    idx = -np.ones((3, n), np.int64)
    for f, (has, other) in enumerate(((col + 1 < grb.ncol, cell + 1),
                                      (rrow + 1 < grb.nrow, cell + ncol),
                                      (lay + 1 < grb.nlay, cell + ncpl))):
        has = has.copy()
        has[has] &= live[has] & live[other[has]]
        q = cell[has] * (n + 1) + other[has]
        pos = np.minimum(np.searchsorted(skey, q), len(skey) - 1)
        if np.any(skey[pos] != q):
            raise Mf6Error("a DIS neighbour is missing from JA")
        idx[f, has] = order[pos]
    return idx


def engine_arrays(grb: Grb, ts: TimeStepData):
    """The engine's inputs of one time step: MODFLOW-2005 face flows frf,
    fff, flf, the 6 boundary flows per cell, IBoundTS, heads, layer types
    (ICELLTYPE) and the source and sink flows."""
    idx = _neighbour_index(grb)
    face = []
    for f in range(3):
        q = np.zeros(grb.ncells)
        on = idx[f] >= 0
        q[on] = -ts.flows_ja[idx[f][on]]
        face.append(q)
    return dict(frf=face[0], fff=face[1], flf=face[2], bflow=ts.boundary_flows.copy(),
                ibound=ts.ibound_ts.copy(), head=ts.heads.copy(),
                layer_type=grb.icelltype.astype(np.int32), source_flow=ts.source_flows.copy(),
                sink_flow=ts.sink_flows.copy())


def problem(grb: Grb, ts: TimeStepData, porosity, retardation=None, **options):
    """The engine's Problem and cell extras (rwpt_set_cell_extras keyword
    arguments) for one read time step.  options: the Problem's dispersion,
    advection, time-step and tracking fields (aL, aT, dmeff, ts_params ...)."""
    from .problems import Problem
    a = engine_arrays(grb, ts)
    n = grb.ncells
    P = Problem(ncol=grb.ncol, nrow=grb.nrow, nlay=grb.nlay, delr=grb.delr.copy(),
                delc=grb.delc.copy(), top=grb.top.copy(), bot=grb.botm.copy(),
                frf=a["frf"], fff=a["fff"], flf=a["flf"], bflow=a["bflow"],
                porosity=np.broadcast_to(np.asarray(porosity, np.float64), (n,)).copy(),
                retardation=None if retardation is None else
                np.broadcast_to(np.asarray(retardation, np.float64), (n,)).copy(),
                ibound=a["ibound"], steady_state=int(bool(ts.steady_state)),
                saturated_top=None if ts.saturated_top is None else ts.saturated_top.copy(),
                **options)
    extras = dict(head=a["head"], layer_type=a["layer_type"], source_flow=a["source_flow"],
                  sink_flow=a["sink_flow"])
    return P, extras


def usg_problem(grb: GrbDisv, ts: TimeStepData, porosity, retardation=None, **options):
    """The engine's UsgProblem for one read time step of a DISV model: the
    grid of ReadGRB/CheckRectangular, FlowModelData's FLOW-JA-FACE,
    boundary, storage, source and sink flows, IBoundTS and heads, ICELLTYPE
    as the layer type (pr_FillTransportCellBufferUnstructured,
    ParticleTrackingEngine.f90:711-748).  options: dispersion, advection,
    time-step and tracking fields (aL, aT, dmeff, ts_params ...)."""
    from .problems import UsgProblem
    n = grb.ncells
    return UsgProblem(
        ncell=n, nlay=grb.nlay, ia=(grb.ia + 1).astype(np.int32),
        ja=grb.ja.astype(np.int32), layer=grb.layer, cell_x=grb.cell_x.copy(),
        cell_y=grb.cell_y.copy(), del_x=grb.del_x.copy(), del_y=grb.del_y.copy(),
        top=grb.top.copy(), bot=grb.botm.copy(), flowja=ts.flows_ja.copy(),
        porosity=np.broadcast_to(np.asarray(porosity, np.float64), (n,)).copy(),
        bflow=ts.boundary_flows.copy(), storage=ts.storage_flows.copy(),
        source=ts.source_flows.copy(), sink=ts.sink_flows.copy(),
        retardation=None if retardation is None else
        np.broadcast_to(np.asarray(retardation, np.float64), (n,)).copy(),
        head=ts.heads.copy(), layer_type=grb.icelltype.astype(np.int32),
        ibound=ts.ibound_ts.copy(), steady_state=int(bool(ts.steady_state)),
        saturated_top=None if ts.saturated_top is None else ts.saturated_top.copy(),
        **options)


class FlowSeries:
    """The flow of successive MODFLOW time steps for Driver(flow_for_step=):
    steps lists (kper, kstp) of the cumulative time steps 1, 2, ...; each
    call reads that step's head and budget records and returns
    problem(...)'s Problem and cell extras."""

    def __init__(self, grb_path, heads_path, budget_path, steps, ibound, hdry, porosity,
                 retardation=None, default_iface=None, **options):
        self.grb = read_grb(grb_path)
        self.heads_path, self.budget_path = heads_path, budget_path
        self.steps = list(steps)
        self.ibound, self.hdry = ibound, hdry
        self.porosity, self.retardation = porosity, retardation
        self.default_iface = default_iface
        self.options = options

    def __call__(self, ktime):
        kper, kstp = self.steps[ktime - 1]
        ts = load_time_step(self.grb, read_heads(self.heads_path, kper, kstp),
                            read_budget(self.budget_path, kper, kstp), kper, kstp,
                            self.ibound, self.hdry, self.default_iface)
        if isinstance(self.grb, GrbDisv):
            # an unstructured problem carries its cell extras
            return usg_problem(self.grb, ts, self.porosity, self.retardation,
                               **self.options), None
        return problem(self.grb, ts, self.porosity, self.retardation, **self.options)

