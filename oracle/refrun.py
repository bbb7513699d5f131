"""Runs the compiled reference driver (oracle/_ref/rwpt_ref, built by
oracle/ref/Makefile from the reference's own Fortran sources) on a problem
and a particle cloud.  Test infrastructure: it checks the oracle, nothing in
the product calls it.

The driver tracks particle k on OpenMP thread k, so the reference's
per-thread rng_par_zig streams become per-particle streams; particle counts
are kept to a few hundred.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BINARY = os.path.join(HERE, "_ref", "rwpt_ref")


def available() -> bool:
    return os.path.exists(BINARY)


def _filled(a, n, value, dtype):
    if a is None:
        return np.full(n, value, dtype)
    return np.ascontiguousarray(a, dtype)


def run(P, cell, lx, ly, lz, t_out, t=None, status=None, timeout=600, threads=None,
        timing=None):
    """Tracks the particles from time 0 through the output times t_out.
    Returns a list (one per output time) of dicts cell/status/lx/ly/lz/t.
    threads: timing mode, the reference's own one rng stream per OpenMP
    thread on that many threads (the draws then differ from the
    per-particle streams); the tracking seconds go to timing["seconds"]."""
    n = int(np.asarray(cell).size)
    nc = P.ncell
    nd = int(np.asarray(P.aL).size)
    t_out = np.ascontiguousarray(t_out, np.float64)
    hdr = np.zeros(20, np.int32)
    hdr[:5] = (P.ncol, P.nrow, P.nlay, n, t_out.size)
    hdr[5], hdr[6] = P.advection_kind, P.time_step_kind
    hdr[7:10] = P.dim_mask
    hdr[10], hdr[11], hdr[12] = P.disp_model, nd, P.default_icbound
    hdr[13], hdr[14] = P.backward, P.uniform_porosity
    hdr[15] = getattr(P, "stop_weak_sinks", 0)
    hdr[16] = getattr(P, "stop_weak_sources", 0)
    hdr[17] = getattr(P, "stop_zone", 0)
    hdr[19] = int(getattr(P, "steady_state", 1))
    f8 = np.float64
    parts = [hdr, np.asarray(P.ts_params, f8),
             _filled(P.delr, P.ncol, 0, f8), _filled(P.delc, P.nrow, 0, f8),
             _filled(P.top, nc, 0, f8), _filled(P.bot, nc, 0, f8),
             _filled(P.frf, nc, 0, f8), _filled(P.fff, nc, 0, f8),
             _filled(P.flf, nc, 0, f8), _filled(P.bflow, 6 * nc, 0, f8),
             _filled(P.porosity, nc, 0, f8), _filled(P.retardation, nc, 1.0, f8),
             # heads at the cell top: the structured fill's layer type is an
             # uninitialised local (ParticleTrackingEngine.f90:801-817), so
             # heads must not matter here
             _filled(P.top, nc, 0, f8),
             _filled(getattr(P, "source_flow", None), nc, 0, f8),
             _filled(getattr(P, "sink_flow", None), nc, 0, f8),
             _filled(P.ibound, nc, 1, np.int32), _filled(P.icbound, nc, 0, np.int32),
             _filled(getattr(P, "zone", None), nc, 1, np.int32),
             _filled(P.aL, nd, 0, f8), _filled(P.aT, nd, 0, f8),
             _filled(P.aLV, nd, 0, f8), _filled(P.aTV, nd, 0, f8),
             _filled(P.dmeff, nd, 0, f8),
             np.ascontiguousarray(cell, np.int32), np.ascontiguousarray(lx, f8),
             np.ascontiguousarray(ly, f8), np.ascontiguousarray(lz, f8),
             _filled(t, n, 0.0, f8), _filled(status, n, 1, np.int32), t_out]
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as fh:
            for a in parts:
                fh.write(np.ascontiguousarray(a).tobytes())
        team = max(n, 1) if threads is None else int(threads)
        env = dict(os.environ, OMP_STACKSIZE="2M", OMP_THREAD_LIMIT=str(team),
                   OMP_DYNAMIC="false")
        args = [BINARY, fin, fout] + ([] if threads is None else [str(team)])
        res = subprocess.run(args, check=True, timeout=timeout, env=env,
                             stdout=subprocess.PIPE, text=True)
        if timing is not None:
            for line in res.stdout.splitlines():
                if line.startswith("seconds"):
                    timing["seconds"] = float(line.split()[1])
        raw = open(fout, "rb").read()
    out, off = [], 0
    for _ in range(t_out.size):
        rec = {}
        for name, dt in (("cell", np.int32), ("status", np.int32), ("lx", f8),
                         ("ly", f8), ("lz", f8), ("t", f8)):
            sz = n * np.dtype(dt).itemsize
            rec[name] = np.frombuffer(raw[off:off + sz], dt).copy()
            off += sz
        out.append(rec)
    return out


GPKDE_BINARY = os.path.join(HERE, "_ref", "gpkde_ref")


def gpkde_available() -> bool:
    return os.access(GPKDE_BINARY, os.X_OK)


def gpkde(x, y, z, origin, bin_size, nbins, weights=None, n_opt_loops=0,
          database=False, isotropic=False, unit_volume=True, scaling_factor=1.0,
          smoothing_selection=0, smoothing_factor=1.0, skip_error=False,
          export_vars=False, bound=0, h_over_delta=(0.0, 0.0, 0.0),
          weight_format=0, export=None, timeout=600, threads=1):
    """Runs the reference's GridProjectedKDE on the points; returns
    (density, histogram counts, initial smoothing) with x fastest, and with
    export_vars also the per-loop optimisation variables (a list of arrays,
    one row per active bin: ix iy iz, density, smoothing(3), shape(3),
    scale, sigma, curvature bandwidth(3), nEstimate, roughness XX YY ZZ,
    net roughness).  bound: boundKernelSizeFormat; h_over_delta: (min,
    delta, max) h/delta passed to Initialize when min > 0; weight_format:
    effectiveWeightFormat; export = (format, column format, time id, time,
    group id) makes the reference write its density rows
    (ExportDensityUnit[Binary]) and appends their bytes to the result.
    threads: OpenMP threads of the reference's loops (1, the default, gives
    its sequential results; timing runs take all host cores)."""
    n = int(np.asarray(x).size)
    ex = export if export is not None else (-1, 0, 0, 0.0, 0)
    hdr = np.array([n, *nbins, n_opt_loops, int(database), int(isotropic),
                    int(weights is not None), int(unit_volume), smoothing_selection,
                    int(export_vars), int(skip_error), bound, weight_format,
                    ex[0], ex[1], ex[2], ex[4]], np.int32)
    f8 = np.float64
    w = np.ones(n) if weights is None else np.ascontiguousarray(weights, f8)
    parts = [hdr, np.asarray(origin, f8), np.asarray(bin_size, f8),
             np.array([scaling_factor, smoothing_factor, *h_over_delta, ex[3]], f8),
             np.ascontiguousarray(x, f8), np.ascontiguousarray(y, f8),
             np.ascontiguousarray(z, f8), w]
    nb = int(np.prod(nbins))
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as fh:
            for a in parts:
                fh.write(np.ascontiguousarray(a).tobytes())
        env = dict(os.environ, OMP_NUM_THREADS=str(max(int(threads), 1)))
        subprocess.run([GPKDE_BINARY, fin, fout], check=True, timeout=timeout,
                       stdout=subprocess.DEVNULL, cwd=d, env=env)
        raw = np.fromfile(fout, f8)
        loops = []
        if export_vars:
            # the reference names the file of loop m gpkde.out<m>: loop 0
            # only without optimisation loops, loops 1.. otherwise
            for k in range(0, n_opt_loops + 1):
                f = os.path.join(d, f"gpkde.out{k}")
                if os.path.exists(f):
                    loops.append(np.loadtxt(f, ndmin=2))
        if export is not None:
            with open(os.path.join(d, "export.out"), "rb") as fh:
                exported = fh.read()
    out = (raw[:nb].copy(), raw[nb:2 * nb].copy(), raw[2 * nb:2 * nb + 3].copy())
    if export_vars:
        out = out + (loops,)
    return out + (exported,) if export is not None else out


WRITERS_BINARY = os.path.join(HERE, "_ref", "writers_ref")


def writers_available() -> bool:
    return os.access(WRITERS_BINARY, os.X_OK)


def writers(P, state, initial, meta, timeout=120):
    """Runs the reference's timeseries, endpoint and observation writers
    (writers_ref) on one grid and particle set; returns the bytes of
    (timeseries, endpoints, mass endpoints, resident records text, binary,
    sink records text, binary).  state: cell, lx, ly, lz, t, status;
    initial: cell, lx, ly, lz, t, face; meta: group, sequence, id, mass,
    solute (per group), names (per group), tracking_direction, time_point,
    time_step, reference_time, origin, rotation, time, zones (per cell) and
    optionally obs_cells (1-based), retardation and water_volume (per cell),
    flow_rate and saturated_top (per cell: the top of ConvertToModelXYZ)."""
    f8, i4 = np.float64, np.int32
    nb = P.ncell
    hdr = np.array([P.ncol, P.nrow, P.nlay, state["cell"].size, len(meta["names"]),
                    meta["tracking_direction"], meta["time_point"], meta["time_step"]], i4)
    names = b"".join(n.encode().ljust(20)[:20] for n in meta["names"])
    parts = [hdr, np.array([meta["reference_time"], *meta["origin"], meta["rotation"],
                            meta["time"]], f8),
             np.asarray(P.delr, f8), np.asarray(P.delc, f8), np.asarray(P.top, f8),
             np.asarray(P.bot, f8), np.asarray(meta["zones"], i4).reshape(nb)]
    for a in (meta["group"], meta["sequence"], meta["id"], state["status"], state["cell"],
              initial["cell"], initial["face"]):
        parts.append(np.ascontiguousarray(a, i4))
    for a in (state["lx"], state["ly"], state["lz"], state["t"], initial["lx"],
              initial["ly"], initial["lz"], initial["t"], meta["mass"]):
        parts.append(np.ascontiguousarray(a, f8))
    parts.append(np.ascontiguousarray(meta["solute"], i4))
    isobs = np.zeros(nb, i4)
    isobs[np.asarray(meta.get("obs_cells", []), np.int64) - 1] = 1
    tail = [isobs, np.asarray(meta.get("retardation", np.ones(nb)), f8).reshape(nb),
            np.asarray(meta.get("water_volume", np.zeros(nb)), f8).reshape(nb),
            np.array([meta.get("flow_rate", 0.0)], f8)]
    if meta.get("saturated_top") is not None:
        tail.append(np.asarray(meta["saturated_top"], f8).reshape(nb))
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        with open(fin, "wb") as fh:
            for a in parts:
                fh.write(np.ascontiguousarray(a).tobytes())
            fh.write(names)
            for a in tail:
                fh.write(np.ascontiguousarray(a).tobytes())
        subprocess.run([WRITERS_BINARY, fin], check=True, timeout=timeout,
                       stdout=subprocess.DEVNULL, cwd=d)
        return tuple(open(os.path.join(d, f), "rb").read()
                     for f in ("ts.out", "ep.out", "epm.out", "obsres.out", "obsres.bin",
                               "obssink.out", "obssink.bin"))


MF6_BINARY = os.path.join(HERE, "_ref", "mf6_ref")


def mf6_available() -> bool:
    return os.access(MF6_BINARY, os.X_OK)


def mf6(grb, hds, cbc, ibound, porosity, retardation, kper=1, kstp=1, hnoflow=1e30,
        hdry=-1e30, default_iface=(), timeout=120, disv=False):
    """Reads MODFLOW 6 DIS grid, head and budget files with the reference's
    readers and FlowModelData%LoadTimeStep (mf6_ref), and assembles every
    cell as ModpathCellData%SetMassTransportDataUnstructured does.  Returns
    FlowModelData's arrays and per cell the first sub-face flows Q1..Q6,
    the cell's source and sink flows and its sub-cell velocities and sizes."""
    ibound = np.ascontiguousarray(ibound, np.int32)
    n = ibound.size
    labels = [k for k, _ in default_iface]
    values = np.array([v for _, v in default_iface], np.int32)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(np.array([kper, kstp], np.int32).tobytes())
            fh.write(np.array([hnoflow, hdry], np.float64).tobytes())
            fh.write(np.array([len(labels)], np.int32).tobytes())
            fh.write(ibound.tobytes())
            fh.write(np.ascontiguousarray(porosity, np.float64).tobytes())
            fh.write(np.ascontiguousarray(retardation, np.float64).tobytes())
            for lab in labels:
                fh.write(lab.encode().ljust(16)[:16])
            fh.write(values.tobytes())
        subprocess.run([MF6_BINARY, str(grb), str(hds), str(cbc), fin, fout] +
                       (["disv"] if disv else []), check=True,
                       timeout=timeout, stdout=subprocess.DEVNULL, cwd=d)
        b = open(fout, "rb").read()
    ncell, nja, steady = np.frombuffer(b[:12], np.int32)
    assert ncell == n
    p = 12
    out = {"steady_state": bool(steady)}
    out["ibound_ts"] = np.frombuffer(b[p:p + 4 * n], np.int32).copy()
    p += 4 * n
    for k in ("heads", "source_flows", "sink_flows", "storage_flows"):
        out[k] = np.frombuffer(b[p:p + 8 * n], np.float64).copy()
        p += 8 * n
    out["boundary_flows"] = np.frombuffer(b[p:p + 48 * n], np.float64).copy()
    p += 48 * n
    out["flows_ja"] = np.frombuffer(b[p:p + 8 * nja], np.float64).copy()
    p += 8 * nja
    for k in ("top", "bottom"):
        out[k] = np.frombuffer(b[p:p + 8 * n], np.float64).copy()
        p += 8 * n
    cells = np.frombuffer(b[p:p + 8 * 17 * n], np.float64).reshape(n, 17)
    out["q"] = cells[:, :6].copy()
    out["cell_source"] = cells[:, 6].copy()
    out["cell_sink"] = cells[:, 7].copy()
    out["velocity"] = cells[:, 8:14].copy()
    out["size"] = cells[:, 14:17].copy()
    p += 8 * 17 * n
    if disv:
        for k in ("cell_x", "cell_y", "del_x", "del_y"):
            out[k] = np.frombuffer(b[p:p + 8 * n], np.float64).copy()
            p += 8 * n
    out["saturated_top"] = np.frombuffer(b[p:p + 8 * n], np.float64).copy()
    return out


USG_BINARY = os.path.join(HERE, "_ref", "usg_ref")


def usg_available() -> bool:
    return os.access(USG_BINARY, os.X_OK)


def _usg_input(U, cell, lx, ly, lz, t_out, t, status):
    n = int(np.asarray(cell).size)
    nc = U.ncell
    nd = int(np.asarray(U.aL).size)
    t_out = np.ascontiguousarray(t_out, np.float64)
    hdr = np.zeros(20, np.int32)
    hdr[0], hdr[1] = n, t_out.size
    hdr[2], hdr[3] = U.advection_kind, U.time_step_kind
    hdr[4:7] = U.dim_mask
    hdr[7], hdr[8], hdr[9] = U.disp_model, nd, U.default_icbound
    hdr[10], hdr[11] = U.backward, U.uniform_porosity
    hdr[12] = U.stop_weak_sinks
    hdr[13] = U.stop_weak_sources
    hdr[14] = U.stop_zone
    hdr[15] = int(U.steady_state)
    f8 = np.float64
    return [hdr, np.asarray(U.ts_params, f8), np.asarray(U.flowja, f8),
            _filled(U.bflow, 6 * nc, 0, f8), _filled(U.storage, nc, 0, f8),
            _filled(U.source, nc, 0, f8), _filled(U.sink, nc, 0, f8),
            _filled(U.porosity, nc, 0, f8), _filled(U.retardation, nc, 1.0, f8),
            _filled(U.head, nc, 0, f8) if U.head is not None else _filled(U.top, nc, 0, f8),
            _filled(U.ibound, nc, 1, np.int32), _filled(U.icbound, nc, 0, np.int32),
            _filled(U.zone, nc, 1, np.int32),
            _filled(U.aL, nd, 0, f8), _filled(U.aT, nd, 0, f8),
            _filled(U.aLV, nd, 0, f8), _filled(U.aTV, nd, 0, f8),
            _filled(U.dmeff, nd, 0, f8),
            np.ascontiguousarray(cell, np.int32), np.ascontiguousarray(lx, f8),
            np.ascontiguousarray(ly, f8), np.ascontiguousarray(lz, f8),
            _filled(t, n, 0.0, f8), _filled(status, n, 1, np.int32), t_out]


def _usg_call(grb, parts, mode, team, timeout):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as fh:
            for a in parts:
                fh.write(np.ascontiguousarray(a).tobytes())
        env = dict(os.environ, OMP_STACKSIZE="2M", OMP_THREAD_LIMIT=str(team),
                   OMP_DYNAMIC="false")
        subprocess.run([USG_BINARY, grb, fin, fout, mode], check=True, timeout=timeout,
                       env=env, stdout=subprocess.PIPE, text=True)
        return open(fout, "rb").read()


def usg_walk(grb, U, cell, lx, ly, lz, t_out, t=None, status=None, timeout=600):
    """Tracks particles on the DISV grid of the binary grid file grb through
    the output times (one OpenMP thread and rng stream per particle).
    Returns a list of dicts cell/status/lx/ly/lz/t, one per output time."""
    n = int(np.asarray(cell).size)
    raw = _usg_call(grb, _usg_input(U, cell, lx, ly, lz, t_out, t, status), "walk",
                    max(n, 1), timeout)
    out, off = [], 0
    for _ in range(np.asarray(t_out).size):
        rec = {}
        for name, dt in (("cell", np.int32), ("status", np.int32), ("lx", np.float64),
                         ("ly", np.float64), ("lz", np.float64), ("t", np.float64)):
            sz = n * np.dtype(dt).itemsize
            rec[name] = np.frombuffer(raw[off:off + sz], dt).copy()
            off += sz
        out.append(rec)
    return out


# fields of one sub-cell record of usg_ref's tables mode
USG_INTS = 18      # row col conn(6) massbnd(6) hasexit cellhasexit dry partiallydry
USG_DOUBLES = 18 + 32 + 8 + 48


def usg_tables(grb, U, timeout=600):
    """Sub-cell tables of every cell: a list (per cell) of lists (per
    sub-cell, row-major) of (ints[18], doubles[106]) as usg_ref writes them:
    ints row col conn(6) massbnd(6) hasexit cellhasexit dry partiallydry;
    doubles VX1..VZ2 DX DY DZ offx(2) offy(2) porosity retardation dtcell
    source sink, qcorner(4 x 8 corners), porosity corners(8), Dxxx Dyyy
    Dzzz Dxyx Dxzx Dyzy (8 each)."""
    z = np.zeros(0)
    raw = _usg_call(grb, _usg_input(U, np.zeros(0, np.int32), z, z, z, np.zeros(1), None,
                                    None), "tables", 1, timeout)
    off = 0
    ncell = int(np.frombuffer(raw[off:off + 4], np.int32)[0])
    off += 4
    cells = []
    for _ in range(ncell):
        nsub = int(np.frombuffer(raw[off:off + 4], np.int32)[0])
        off += 4
        subs = []
        for _ in range(nsub):
            ints = np.frombuffer(raw[off:off + 4 * USG_INTS], np.int32).copy()
            off += 4 * USG_INTS
            dbl = np.frombuffer(raw[off:off + 8 * USG_DOUBLES], np.float64).copy()
            off += 8 * USG_DOUBLES
            subs.append((ints, dbl))
        cells.append(subs)
    return cells


MPATHRW_BINARY = os.path.join(HERE, "_ref", "mpathrw")


def mpathrw_available() -> bool:
    return os.access(MPATHRW_BINARY, os.X_OK)


def mpathrw(sim_path, timeout=120):
    """Runs MODPATH-RW itself (MPathRW.F90 compiled from the reference's
    sources) on a simulation file, serially, in the file's folder; returns
    its console output.  Raises unless it ends with "Normal termination"."""
    d, name = os.path.split(os.path.abspath(sim_path))
    r = subprocess.run([MPATHRW_BINARY, "-nl", name], cwd=d, stdin=subprocess.DEVNULL,
                       capture_output=True, text=True, timeout=timeout,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    if r.returncode != 0 or "Normal termination" not in r.stdout:
        raise RuntimeError(f"mpathrw failed ({r.returncode}): {r.stdout[-2000:]}"
                           f"{r.stderr[-2000:]}")
    return r.stdout
