"""Writers of small MODFLOW 6 DIS output files for the reader tests.

No MODFLOW 6 output ships with the reference and MODFLOW 6 is not in this
image, so the tests write the three files themselves in MODFLOW 6's binary
layouts (stream access, little-endian): the binary grid file (four
50-character header lines, 16 definition lines of 100 characters, the
variables), structured double-precision head records, and compact budget
records (method 1 arrays, method 6 lists with model and package names and
auxiliary values).  Both the reader under test and the reference's own
readers (oracle/_ref/mf6_ref) read the same files.
"""
import numpy as np

_DEFS = [
    ("NCELLS", "INTEGER NDIM 0 # NCELLS"), ("NLAY", "INTEGER NDIM 0 # NLAY"),
    ("NROW", "INTEGER NDIM 0 # NROW"), ("NCOL", "INTEGER NDIM 0 # NCOL"),
    ("NJA", "INTEGER NDIM 0 # NJA"), ("XORIGIN", "DOUBLE NDIM 0 # XORIGIN"),
    ("YORIGIN", "DOUBLE NDIM 0 # YORIGIN"), ("ANGROT", "DOUBLE NDIM 0 # ANGROT"),
    ("DELR", "DOUBLE NDIM 1 NCOL"), ("DELC", "DOUBLE NDIM 1 NROW"),
    ("TOP", "DOUBLE NDIM 1 NCPL"), ("BOTM", "DOUBLE NDIM 1 NCELLS"),
    ("IA", "INTEGER NDIM 1 NCELLS+1"), ("JA", "INTEGER NDIM 1 NJA"),
    ("IDOMAIN", "INTEGER NDIM 1 NCELLS"), ("ICELLTYPE", "INTEGER NDIM 1 NCELLS"),
]


def _txt(s, n):
    b = s.encode("ascii")
    assert len(b) <= n
    return b + b" " * (n - len(b))


def connections(nlay, nrow, ncol, idomain=None):
    """IA (1-based) and JA of a DIS grid in MODFLOW 6's user numbering: each
    cell first, then its neighbours in increasing cell number.  A cell with
    IDOMAIN < 1 is removed: its row is empty and no row lists it (the
    binary grid file of a DIS model with removed cells maps the reduced
    connections to user cell numbers this way, so FLOW-JA-FACE keeps JA's
    positions)."""
    ncpl = nrow * ncol
    n = nlay * nrow * ncol
    on = np.ones(n, bool) if idomain is None else np.asarray(idomain) >= 1
    ia, ja = [1], []
    for c in range(n):
        if on[c]:
            k, r, j = c // ncpl, (c % ncpl) // ncol, c % ncol
            nb = []
            if k > 0: nb.append(c - ncpl)
            if r > 0: nb.append(c - ncol)
            if j > 0: nb.append(c - 1)
            if j + 1 < ncol: nb.append(c + 1)
            if r + 1 < nrow: nb.append(c + ncol)
            if k + 1 < nlay: nb.append(c + ncpl)
            ja += [c + 1] + [m + 1 for m in sorted(nb) if on[m]]
        ia.append(len(ja) + 1)
    return np.array(ia, np.int32), np.array(ja, np.int32)


def write_grb(path, nlay, nrow, ncol, delr, delc, top, botm, icelltype, idomain=None):
    n = nlay * nrow * ncol
    idomain = np.ones(n, np.int32) if idomain is None else idomain
    ia, ja = connections(nlay, nrow, ncol, idomain)
    with open(path, "wb") as f:
        for line in ("GRID DIS", "VERSION 1", "NTXT 16", "LENTXT 100"):
            f.write(_txt(line, 50))
        for name, rest in _DEFS:
            f.write(_txt(f"{name} {rest}", 100))
        f.write(np.array([n, nlay, nrow, ncol, len(ja)], "<i4").tobytes())
        f.write(np.array([0.0, 0.0, 0.0], "<f8").tobytes())
        for a in (delr, delc, top, botm):
            f.write(np.asarray(a, "<f8").tobytes())
        f.write((ia).astype("<i4").tobytes())
        f.write(ja.astype("<i4").tobytes())
        f.write(np.asarray(idomain, "<i4").tobytes())
        f.write(np.asarray(icelltype, "<i4").tobytes())
    return ia, ja


def write_heads(path, nlay, nrow, ncol, head, steps=((1, 1, 1.0),)):
    """One record per layer and time step (kper, kstp, totim); head is one
    array for every step or a list of one array per step."""
    ncpl = nrow * ncol
    with open(path, "wb") as f:
        for si, (kper, kstp, totim) in enumerate(steps):
            h = head[si] if isinstance(head, (list, tuple)) else head
            for k in range(nlay):
                f.write(np.array([kstp, kper], "<i4").tobytes())
                f.write(np.array([totim, totim], "<f8").tobytes())
                f.write(_txt("HEAD", 16))
                f.write(np.array([ncol, nrow, k + 1], "<i4").tobytes())
                f.write(np.asarray(h[k * ncpl:(k + 1) * ncpl], "<f8").tobytes())


def flow_ja(ia, ja, frf, fff, flf, nrow, ncol):
    """FLOW-JA-FACE from MODFLOW-2005 face flows: F(n, m) is the flow into n
    from m, antisymmetric."""
    ncpl = nrow * ncol
    out = np.zeros(len(ja))
    for c in range(len(ia) - 1):
        for p in range(ia[c] - 1, ia[c + 1] - 1):
            m = ja[p] - 1
            if m == c:
                continue
            if m == c + 1:
                out[p] = -frf[c]
            elif m == c - 1:
                out[p] = frf[m]
            elif m == c + ncol:
                out[p] = -fff[c]
            elif m == c - ncol:
                out[p] = fff[m]
            elif m == c + ncpl:
                out[p] = -flf[c]
            else:
                out[p] = flf[m]
    return out


def _head(f, kper, kstp, text, n1, n2, n3, method, totim):
    f.write(np.array([kstp, kper], "<i4").tobytes())
    f.write(_txt(text.rjust(16), 16))
    f.write(np.array([n1, n2, -n3, method], "<i4").tobytes())
    f.write(np.array([1.0, totim, totim], "<f8").tobytes())


def write_budget(path, nlay, nrow, ncol, fja, records, kper=1, kstp=1, totim=1.0,
                 append=False):
    """FLOW-JA-FACE, then each record: ("array", text, values) written as a
    method 1 array over the grid, or ("list", text, package, nodes, q,
    aux_names, aux) written as a method 6 list."""
    with open(path, "ab" if append else "wb") as f:
        _head(f, kper, kstp, "FLOW-JA-FACE", len(fja), 1, 1, 1, totim)
        f.write(np.asarray(fja, "<f8").tobytes())
        for r in records:
            if r[0] == "array":
                _, text, values = r
                _head(f, kper, kstp, text, ncol, nrow, nlay, 1, totim)
                f.write(np.asarray(values, "<f8").tobytes())
            else:
                _, text, pkg, nodes, q, aux_names, aux = r
                _head(f, kper, kstp, text, ncol, nrow, nlay, 6, totim)
                for s in ("MODEL", "MODEL", "MODEL", pkg):
                    f.write(_txt(s, 16))
                f.write(np.array([1 + len(aux_names)], "<i4").tobytes())
                for a in aux_names:
                    f.write(_txt(a, 16))
                f.write(np.array([len(nodes)], "<i4").tobytes())
                for i, (nd, v) in enumerate(zip(nodes, q)):
                    f.write(np.array([nd, i + 1], "<i4").tobytes())
                    f.write(np.array([v] + [float(x) for x in (aux[i] if aux_names else [])],
                                     "<f8").tobytes())
