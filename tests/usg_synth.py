"""Synthetic MODFLOW 6 DISV grids of rectangular cells (smoothed quadtrees)
for the unstructured-grid tests.

A base grid of nrow x ncol cells (row 0 at the top, the largest y) has some
cells split into 2 x 2 children; every horizontal face then meets either one
cell of its size, one cell twice its size or two cells of half its size, the
"smoothed rectangular" grids RectangularGridDisvMf6.f90:326-570 accepts.
Cells are numbered layer by layer; within a layer base rows top to bottom,
columns left to right, a refined base cell contributing its children
upper-left, upper-right, lower-left, lower-right.

write_grb writes the binary grid file in MODFLOW 6's DISV layout (the
variables the reference's ReadGRB reads, :47-274): vertices listed
clockwise from the upper-left corner, closed, with a mid-face vertex on the
sides of a cell that meet two smaller cells.

Flows: horizontal FLOW-JA-FACE from a stream function psi evaluated at the
face end points (the flow across a segment is the difference of psi at its
ends times the layer thickness, so every cell's horizontal flows balance up
to rounding), vertical flows from a seeded field.  FLOW-JA-FACE(n, m) is
the flow into n from m.
"""
import numpy as np

from mf6_synth import _txt


class Quadtree:
    def __init__(self, nrow, ncol, refine, delr=1.0, delc=1.0, nlay=1, top=None,
                 botm=None):
        self.nrow, self.ncol, self.nlay = nrow, ncol, nlay
        self.refine = set(refine)
        cells = []  # (x0, x1, y0, y1, base_i, base_j)
        ytop = nrow * delc
        for i in range(nrow):
            for j in range(ncol):
                x0, x1 = j * delr, (j + 1) * delr
                y1, y0 = ytop - i * delc, ytop - (i + 1) * delc
                if (i, j) in self.refine:
                    xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
                    cells += [(x0, xm, ym, y1, i, j), (xm, x1, ym, y1, i, j),
                              (x0, xm, y0, ym, i, j), (xm, x1, y0, ym, i, j)]
                else:
                    cells.append((x0, x1, y0, y1, i, j))
        self.layer_cells = cells
        self.ncpl = len(cells)
        self.ncell = self.ncpl * nlay
        b = np.array([c[:4] for c in cells], float)
        self.x0, self.x1, self.y0, self.y1 = b[:, 0], b[:, 1], b[:, 2], b[:, 3]
        self.cell_x = 0.5 * (self.x0 + self.x1)
        self.cell_y = 0.5 * (self.y0 + self.y1)
        self.del_x = self.x1 - self.x0
        self.del_y = self.y1 - self.y0
        if top is None:
            top = np.full(self.ncpl, float(nlay))
        if botm is None:
            botm = np.concatenate([np.full(self.ncpl, float(nlay - 1 - k))
                                   for k in range(nlay)])
        self.top_cpl = np.asarray(top, float)
        self.botm = np.asarray(botm, float)
        self.top = np.concatenate([self.top_cpl] + [self.botm[(k - 1) * self.ncpl:
                                                               k * self.ncpl]
                                                    for k in range(1, nlay)])
        self._connect()

    # horizontal neighbours within a layer: rectangles that share a side of
    # positive length
    def _hneighbours(self):
        nb = [[] for _ in range(self.ncpl)]
        for a in range(self.ncpl):
            for c in range(a + 1, self.ncpl):
                xo = min(self.x1[a], self.x1[c]) - max(self.x0[a], self.x0[c])
                yo = min(self.y1[a], self.y1[c]) - max(self.y0[a], self.y0[c])
                touch_x = (self.x1[a] == self.x0[c] or self.x1[c] == self.x0[a]) and yo > 0
                touch_y = (self.y1[a] == self.y0[c] or self.y1[c] == self.y0[a]) and xo > 0
                if touch_x or touch_y:
                    nb[a].append(c)
                    nb[c].append(a)
        return nb

    def _connect(self):
        hnb = self._hneighbours()
        ia, ja = [1], []
        for n in range(self.ncell):
            k, p = divmod(n, self.ncpl)
            nb = [k * self.ncpl + c for c in hnb[p]]
            if k > 0:
                nb.append(n - self.ncpl)
            if k + 1 < self.nlay:
                nb.append(n + self.ncpl)
            ja += [n + 1] + [m + 1 for m in sorted(nb)]
            ia.append(len(ja) + 1)
        self.ia = np.array(ia, np.int32)
        self.ja = np.array(ja, np.int32)
        self.layer = np.repeat(np.arange(1, self.nlay + 1, dtype=np.int32), self.ncpl)

    # ------------------------------------------------------------------
    def vertices(self):
        """Vertex list and per-cell closed clockwise vertex loops."""
        index = {}
        verts = []

        def vid(x, y):
            key = (float(x), float(y))
            if key not in index:
                index[key] = len(verts) + 1
                verts.append(key)
            return index[key]

        loops = []
        for p in range(self.ncpl):
            x0, x1, y0, y1 = self.x0[p], self.x1[p], self.y0[p], self.y1[p]
            xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
            i, j = self.layer_cells[p][4], self.layer_cells[p][5]
            small = (i, j) in self.refine

            def split(di, dj):
                return (not small) and (i + di, j + dj) in self.refine

            loop = [vid(x0, y1)]
            if split(-1, 0):
                loop.append(vid(xm, y1))
            loop.append(vid(x1, y1))
            if split(0, 1):
                loop.append(vid(x1, ym))
            loop.append(vid(x1, y0))
            if split(1, 0):
                loop.append(vid(xm, y0))
            loop.append(vid(x0, y0))
            if split(0, -1):
                loop.append(vid(x0, ym))
            loop.append(loop[0])
            loops.append(loop)
        return np.array(verts, float), loops

    def psi(self, x, y, q0=0.5, amp=0.15):
        """The stream function of flows(): a uniform flow towards -x plus a
        cellular perturbation."""
        Lx = self.ncol * float(self.del_x.max())
        Ly = self.nrow * float(self.del_y.max())
        return -q0 * y + amp * np.sin(2 * np.pi * x / Lx) * np.sin(np.pi * y / Ly) * Ly

    def boundary_flows(self, q0=0.5, amp=0.15):
        """The flows of flows()' field across the west and east model faces:
        (cells, face, flow into the cell) of the cells on x = 0 (face 1) and
        on the east edge (face 2)."""
        Lx = self.ncol * float(self.del_x.max())
        thick = self.top - self.botm
        out = []
        for n in range(self.ncell):
            p = n % self.ncpl
            for x, face, sign in ((0.0, 1, 1.0), (Lx, 2, -1.0)):
                if (face == 1 and self.x0[p] == 0.0) or (face == 2 and self.x1[p] == Lx):
                    fx = (self.psi(x, self.y1[p], q0, amp) -
                          self.psi(x, self.y0[p], q0, amp)) * thick[n]
                    out.append((n + 1, face, sign * fx))
        return out

    def flows(self, seed=0, q0=0.5, amp=0.15, wscale=0.02):
        """FLOW-JA-FACE of a stream-function horizontal field plus seeded
        vertical flows."""
        rng = np.random.default_rng(seed)

        def psi(x, y):
            return self.psi(x, y, q0, amp)

        thick = self.top - self.botm
        w = rng.normal(0.0, wscale, size=(max(self.nlay - 1, 0), self.ncpl))
        out = np.zeros(len(self.ja))
        for n in range(self.ncell):
            k, p = divmod(n, self.ncpl)
            for q in range(self.ia[n], self.ia[n + 1] - 1):
                m = self.ja[q] - 1
                if m == n:
                    continue
                km, pm = divmod(m, self.ncpl)
                if km != k:
                    area = self.del_x[p] * self.del_y[p]
                    if km > k:   # m below n: flow into n from below
                        out[q] = w[k, p] * area
                    else:
                        out[q] = -w[km, p] * area
                    continue
                # shared segment and the outward direction from n
                if self.x1[p] == self.x0[pm] or self.x0[p] == self.x1[pm]:
                    xs = self.x1[p] if self.x1[p] == self.x0[pm] else self.x0[p]
                    ya = max(self.y0[p], self.y0[pm])
                    yb = min(self.y1[p], self.y1[pm])
                    fx = (psi(xs, yb) - psi(xs, ya)) * thick[n]  # flux in +x
                    out[q] = -fx if self.x1[p] == self.x0[pm] else fx
                else:
                    ys = self.y1[p] if self.y1[p] == self.y0[pm] else self.y0[p]
                    xa = max(self.x0[p], self.x0[pm])
                    xb = min(self.x1[p], self.x1[pm])
                    fy = (psi(xa, ys) - psi(xb, ys)) * thick[n]  # flux in +y
                    out[q] = -fy if self.y1[p] == self.y0[pm] else fy
        return out

    def write_grb(self, path, icelltype=None, idomain=None):
        verts, loops = self.vertices()
        javert = np.concatenate([np.array(l, np.int32) for l in loops])
        iavert = np.cumsum([1] + [len(l) for l in loops]).astype(np.int32)
        icelltype = np.zeros(self.ncell, np.int32) if icelltype is None else icelltype
        idomain = np.ones(self.ncell, np.int32) if idomain is None else idomain
        defs = [
            ("NCELLS", "INTEGER NDIM 0 # NCELLS"), ("NLAY", "INTEGER NDIM 0 # NLAY"),
            ("NCPL", "INTEGER NDIM 0 # NCPL"), ("NVERT", "INTEGER NDIM 0 # NVERT"),
            ("NJAVERT", "INTEGER NDIM 0 # NJAVERT"), ("NJA", "INTEGER NDIM 0 # NJA"),
            ("XORIGIN", "DOUBLE NDIM 0 # XORIGIN"), ("YORIGIN", "DOUBLE NDIM 0 # YORIGIN"),
            ("ANGROT", "DOUBLE NDIM 0 # ANGROT"), ("TOP", "DOUBLE NDIM 1 NCPL"),
            ("BOTM", "DOUBLE NDIM 1 NCELLS"), ("VERTICES", "DOUBLE NDIM 2 2 NVERT"),
            ("CELLX", "DOUBLE NDIM 1 NCPL"), ("CELLY", "DOUBLE NDIM 1 NCPL"),
            ("IAVERT", "INTEGER NDIM 1 NCPL+1"), ("JAVERT", "INTEGER NDIM 1 NJAVERT"),
            ("IA", "INTEGER NDIM 1 NCELLS+1"), ("JA", "INTEGER NDIM 1 NJA"),
            ("IDOMAIN", "INTEGER NDIM 1 NCELLS"), ("ICELLTYPE", "INTEGER NDIM 1 NCELLS"),
        ]
        with open(path, "wb") as f:
            for line in ("GRID DISV", "VERSION 1", f"NTXT {len(defs)}", "LENTXT 100"):
                f.write(_txt(line, 50))
            for name, rest in defs:
                f.write(_txt(f"{name} {rest}", 100))
            f.write(np.array([self.ncell, self.nlay, self.ncpl, len(verts), len(javert),
                              len(self.ja)], "<i4").tobytes())
            f.write(np.array([0.0, 0.0, 0.0], "<f8").tobytes())
            f.write(self.top_cpl.astype("<f8").tobytes())
            f.write(self.botm.astype("<f8").tobytes())
            f.write(verts.astype("<f8").tobytes())
            f.write(self.cell_x.astype("<f8").tobytes())
            f.write(self.cell_y.astype("<f8").tobytes())
            f.write(iavert.astype("<i4").tobytes())
            f.write(javert.astype("<i4").tobytes())
            f.write(self.ia.astype("<i4").tobytes())
            f.write(self.ja.astype("<i4").tobytes())
            f.write(np.asarray(idomain, "<i4").tobytes())
            f.write(np.asarray(icelltype, "<i4").tobytes())

    # per-cell arrays over all layers (CellX/CellY/DelX/DelY repeat per layer)
    def cell_arrays(self):
        rep = lambda a: np.tile(a, self.nlay)
        return dict(cell_x=rep(self.cell_x), cell_y=rep(self.cell_y),
                    del_x=rep(self.del_x), del_y=rep(self.del_y),
                    top=self.top.copy(), bot=self.botm.copy(), layer=self.layer.copy())
