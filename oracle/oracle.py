"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this.
The oracle restates the reference algorithm on the CPU; see rwpt_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)


class OrcProblem(C.Structure):
    _fields_ = [
        ("ncol", C.c_int), ("nrow", C.c_int), ("nlay", C.c_int),
        ("delr", _dp), ("delc", _dp), ("top", _dp), ("bot", _dp),
        ("ibound", _ip), ("frf", _dp), ("fff", _dp), ("flf", _dp),
        ("bflow", _dp), ("porosity", _dp), ("retardation", _dp),
        ("head", _dp), ("layer_type", _ip), ("source_flow", _dp),
        ("sink_flow", _dp), ("zone", _ip), ("icbound", _ip),
        ("default_icbound", C.c_int), ("disp_model", C.c_int),
        ("disp_uniform", C.c_int),
        ("aL", _dp), ("aT", _dp), ("aLV", _dp), ("aTV", _dp), ("dmeff", _dp),
        ("advection_kind", C.c_int), ("time_step_kind", C.c_int),
        ("ts_params", C.c_double * 2), ("dim_mask", C.c_int * 3),
        ("uniform_porosity", C.c_int), ("backward", C.c_int),
        ("stop_weak_sinks", C.c_int), ("stop_weak_sources", C.c_int),
        ("stop_zone", C.c_int), ("steady_state", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run build())")
        L = C.CDLL(path)
        L.orc_advance.argtypes = [C.POINTER(OrcProblem), C.c_long, _ip, _dp, _dp,
                                  _dp, _dp, _ip, _u64p, C.c_double, C.c_double,
                                  C.POINTER(C.c_longlong), C.c_int]
        L.orc_advance.restype = C.c_int
        L.orc_cell_tables.argtypes = [C.POINTER(OrcProblem), _dp]
        L.orc_rng_streams.argtypes = [_u64p, C.c_long, C.c_long, _u64p]
        L.orc_xoroshiro_next.argtypes = [_u64p]
        L.orc_xoroshiro_next.restype = C.c_uint64
        L.orc_uniform.argtypes = [_u64p]
        L.orc_uniform.restype = C.c_double
        L.orc_normal12.argtypes = [_u64p]
        L.orc_normal12.restype = C.c_double
        L.orc_rng_jump.argtypes = [_u64p]
        L.orc_global_xyz.argtypes = [C.POINTER(OrcProblem), C.c_long, _ip, _dp,
                                     _dp, _dp, _dp, _dp, _dp]
        L.orc_histogram.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, _dp,
                                    _ip, _dp]
        L.orc_kde_density.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                      _dp, _dp, _dp]
        L.orc_kde_density.restype = C.c_int
        L.orc_kde_density_sel.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, _dp,
                                          _ip, C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_kde_density_sel.restype = C.c_int
        L.orc_newton_axis.argtypes = [C.c_double] * 8
        L.orc_newton_axis.restype = C.c_double
        L.orc_quadratic_axis.argtypes = [C.c_double] * 7 + [_ip]
        L.orc_quadratic_axis.restype = C.c_double
        _LIB = L
    return _LIB


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class OracleProblem:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, P):
        self._keep = {}
        s = OrcProblem()
        s.ncol, s.nrow, s.nlay = P.ncol, P.nrow, P.nlay

        def dd(name, arr):
            a = _d(arr)
            self._keep[name] = a
            return _ptr(a, _dp)

        def ii(name, arr):
            a = _i(arr)
            self._keep[name] = a
            return _ptr(a, _ip)

        s.delr = dd("delr", P.delr)
        s.delc = dd("delc", P.delc)
        s.top = dd("top", P.top)
        s.bot = dd("bot", P.bot)
        s.ibound = ii("ibound", P.ibound)
        s.frf = dd("frf", P.frf)
        s.fff = dd("fff", P.fff)
        s.flf = dd("flf", P.flf)
        s.bflow = dd("bflow", P.bflow)
        s.porosity = dd("porosity", P.porosity)
        s.retardation = dd("retardation", P.retardation)
        s.head = dd("head", getattr(P, "head", None))
        s.layer_type = ii("layer_type", getattr(P, "layer_type", None))
        s.source_flow = dd("source_flow", getattr(P, "source_flow", None))
        s.sink_flow = dd("sink_flow", getattr(P, "sink_flow", None))
        s.zone = ii("zone", getattr(P, "zone", None))
        s.icbound = ii("icbound", P.icbound)
        s.default_icbound = P.default_icbound
        s.disp_model = P.disp_model
        s.disp_uniform = P.disp_uniform
        s.aL = dd("aL", P.aL)
        s.aT = dd("aT", P.aT)
        s.aLV = dd("aLV", P.aLV)
        s.aTV = dd("aTV", P.aTV)
        s.dmeff = dd("dmeff", P.dmeff)
        s.advection_kind = P.advection_kind
        s.time_step_kind = P.time_step_kind
        s.ts_params[0], s.ts_params[1] = P.ts_params
        for d in range(3):
            s.dim_mask[d] = P.dim_mask[d]
        s.uniform_porosity = P.uniform_porosity
        s.backward = P.backward
        s.stop_weak_sinks = int(getattr(P, "stop_weak_sinks", 0))
        s.stop_weak_sources = int(getattr(P, "stop_weak_sources", 0))
        s.stop_zone = int(getattr(P, "stop_zone", 0))
        s.steady_state = int(getattr(P, "steady_state", 1))
        self.s = s


def rng_streams(n, first=0, base=None):
    from_base = np.array(base if base is not None else
                         (0x1DADBEEFBAADD0D0, 0x5BADD0D0DEADBEEF), dtype=np.uint64)
    out = np.zeros(2 * n, dtype=np.uint64)
    lib().orc_rng_streams(from_base.ctypes.data_as(_u64p), first, n,
                          out.ctypes.data_as(_u64p))
    return out


def advance(P, state, t_engine, t_stop, nthreads=0):
    """state: dict with cell, lx, ly, lz, t, status, rng, substeps (modified
    in place).  Returns the oracle's return code."""
    OP = P if isinstance(P, OracleProblem) else OracleProblem(P)
    n = state["cell"].size
    rc = lib().orc_advance(C.byref(OP.s), n,
                           _ptr(state["cell"], _ip), _ptr(state["lx"], _dp),
                           _ptr(state["ly"], _dp), _ptr(state["lz"], _dp),
                           _ptr(state["t"], _dp), _ptr(state["status"], _ip),
                           _ptr(state["rng"], _u64p), t_engine, t_stop,
                           state["substeps"].ctypes.data_as(C.POINTER(C.c_longlong)),
                           nthreads)
    return rc


def cell_tables(P):
    OP = P if isinstance(P, OracleProblem) else OracleProblem(P)
    out = np.zeros((P.ncell if not isinstance(P, OracleProblem) else
                    OP.s.ncol * OP.s.nrow * OP.s.nlay, 104))
    lib().orc_cell_tables(C.byref(OP.s), out.ctypes.data_as(_dp))
    return out


def _with_saturated_top(P):
    # ConvertToModelXYZ's top is Grid%SaturatedTop (ModflowRectangularGrid.f90:
    # 524-549); a problem without one uses the cell tops
    sat = getattr(P, "saturated_top", None)
    return P if sat is None else dataclasses.replace(P, top=np.asarray(sat, np.float64))


def global_xyz(P, cell, lx, ly, lz):
    OP = P if isinstance(P, OracleProblem) else OracleProblem(_with_saturated_top(P))
    n = cell.size
    gx, gy, gz = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_global_xyz(C.byref(OP.s), n, _ptr(_i(cell), _ip), _ptr(_d(lx), _dp),
                         _ptr(_d(ly), _dp), _ptr(_d(lz), _dp), _ptr(gx, _dp),
                         _ptr(gy, _dp), _ptr(gz, _dp))
    return gx, gy, gz


def histogram(x, y, z, origin, binsize, nbins, w=None):
    n = x.size
    counts = np.zeros(int(np.prod(nbins)))
    o = _d(origin)
    b = _d(binsize)
    nb = _i(nbins)
    ww = _d(w)
    lib().orc_histogram(n, _ptr(_d(x), _dp), _ptr(_d(y), _dp), _ptr(_d(z), _dp),
                        _ptr(ww, _dp), _ptr(o, _dp), _ptr(b, _dp), _ptr(nb, _ip),
                        _ptr(counts, _dp))
    return counts


def new_state(cell, lx, ly, lz, t0=0.0, rng=None):
    n = cell.size
    return {
        "cell": _i(cell).copy(), "lx": _d(lx).copy(), "ly": _d(ly).copy(),
        "lz": _d(lz).copy(), "t": np.full(n, t0), "status": np.ones(n, np.int32),
        "rng": rng_streams(n) if rng is None else rng.copy(),
        "substeps": np.zeros(n, np.int64),
    }


def kde_density(x, y, z, origin, binsize, nbins, w=None, selection=0, factor=1.0):
    """gpkde density (nOptimizationLoops = 0, brute-force kernel); selection
    and factor choose the initial smoothing."""
    n = x.size
    nb = int(np.prod(nbins))
    dens = np.zeros(nb)
    counts = np.zeros(nb)
    h = np.zeros(3)
    L = lib()
    L.orc_kde_density_sel(n, _ptr(_d(x), _dp), _ptr(_d(y), _dp), _ptr(_d(z), _dp),
                          _ptr(_d(w), _dp), _ptr(_d(origin), _dp),
                          _ptr(_d(binsize), _dp), _ptr(_i(nbins), _ip),
                          C.c_int(selection), C.c_double(factor), _ptr(dens, _dp),
                          _ptr(counts, _dp), _ptr(h, _dp))
    return dens, counts, h


# --------------------------------------------------------------------------
# Unstructured grids (rwpt_oracle_usg.c)
# --------------------------------------------------------------------------

class OrcUsgProblem(C.Structure):
    _fields_ = [
        ("base", OrcProblem), ("ncell", C.c_int), ("nlay", C.c_int),
        ("nja", C.c_long), ("ia", _ip), ("ja", _ip), ("layer", _ip),
        ("cell_x", _dp), ("cell_y", _dp), ("del_x", _dp), ("del_y", _dp),
        ("flowja", _dp), ("storage", _dp),
    ]


USG_FIELDS = 126


def _usg_lib():
    L = lib()
    if not hasattr(L, "_usg_ready"):
        L.orc_usg_tables.argtypes = [C.POINTER(OrcUsgProblem), _dp]
        L.orc_usg_tables.restype = C.c_long
        L.orc_usg_advance.argtypes = [C.POINTER(OrcUsgProblem), C.c_long, _ip, _dp, _dp,
                                      _dp, _dp, _ip, _u64p, C.c_double, C.c_double,
                                      C.POINTER(C.c_longlong), C.c_int]
        L.orc_usg_advance.restype = C.c_int
        L.orc_usg_global_xyz.argtypes = [C.POINTER(OrcUsgProblem), C.c_long, _ip, _dp,
                                         _dp, _dp, _dp, _dp, _dp]
        L._usg_ready = True
    return L


class OracleUsgProblem:
    """The C view of an unstructured problem U (attributes as
    rwpt.problems.UsgProblem); keeps the arrays alive."""

    def __init__(self, U):
        self._keep = {}
        s = OrcUsgProblem()
        b = s.base

        def dd(name, arr):
            a = _d(arr)
            self._keep[name] = a
            return _ptr(a, _dp)

        def ii(name, arr):
            a = _i(arr)
            self._keep[name] = a
            return _ptr(a, _ip)

        b.top = dd("top", U.top)
        b.bot = dd("bot", U.bot)
        b.ibound = ii("ibound", U.ibound)
        b.bflow = dd("bflow", U.bflow)
        b.porosity = dd("porosity", U.porosity)
        b.retardation = dd("retardation", U.retardation)
        b.head = dd("head", U.head)
        b.layer_type = ii("layer_type", U.layer_type)
        b.source_flow = dd("source_flow", U.source)
        b.sink_flow = dd("sink_flow", U.sink)
        b.zone = ii("zone", U.zone)
        b.icbound = ii("icbound", U.icbound)
        b.default_icbound = U.default_icbound
        b.disp_model = U.disp_model
        b.disp_uniform = 1 if np.asarray(U.aL).size == 1 else 0
        b.aL = dd("aL", U.aL)
        b.aT = dd("aT", U.aT)
        b.aLV = dd("aLV", U.aLV)
        b.aTV = dd("aTV", U.aTV)
        b.dmeff = dd("dmeff", U.dmeff)
        b.advection_kind = U.advection_kind
        b.time_step_kind = U.time_step_kind
        b.ts_params[0], b.ts_params[1] = U.ts_params
        for d in range(3):
            b.dim_mask[d] = U.dim_mask[d]
        b.uniform_porosity = U.uniform_porosity
        b.backward = U.backward
        b.stop_weak_sinks = int(U.stop_weak_sinks)
        b.stop_weak_sources = int(U.stop_weak_sources)
        b.stop_zone = int(U.stop_zone)
        b.steady_state = int(U.steady_state)
        s.ncell, s.nlay, s.nja = U.ncell, U.nlay, len(U.ja)
        s.ia = ii("ia", U.ia)
        s.ja = ii("ja", U.ja)
        s.layer = ii("layer", U.layer)
        s.cell_x = dd("cell_x", U.cell_x)
        s.cell_y = dd("cell_y", U.cell_y)
        s.del_x = dd("del_x", U.del_x)
        s.del_y = dd("del_y", U.del_y)
        s.flowja = dd("flowja", U.flowja)
        s.storage = dd("storage", U.storage)
        self.s = s


def usg_tables(U):
    """(n_subcells, USG_FIELDS) sub-cell tables (layout in rwpt_oracle_usg.c)."""
    OP = U if isinstance(U, OracleUsgProblem) else OracleUsgProblem(U)
    L = _usg_lib()
    k = L.orc_usg_tables(C.byref(OP.s), None)
    if k < 0:
        raise RuntimeError(f"orc_usg_tables: error {k}")
    out = np.zeros((k, USG_FIELDS))
    L.orc_usg_tables(C.byref(OP.s), out.ctypes.data_as(_dp))
    return out


def usg_advance(U, state, t_engine, t_stop, nthreads=0):
    OP = U if isinstance(U, OracleUsgProblem) else OracleUsgProblem(U)
    n = state["cell"].size
    return _usg_lib().orc_usg_advance(
        C.byref(OP.s), n, _ptr(state["cell"], _ip), _ptr(state["lx"], _dp),
        _ptr(state["ly"], _dp), _ptr(state["lz"], _dp), _ptr(state["t"], _dp),
        _ptr(state["status"], _ip), _ptr(state["rng"], _u64p), t_engine, t_stop,
        state["substeps"].ctypes.data_as(C.POINTER(C.c_longlong)), nthreads)


def usg_global_xyz(U, cell, lx, ly, lz):
    OP = U if isinstance(U, OracleUsgProblem) else OracleUsgProblem(_with_saturated_top(U))
    n = cell.size
    gx, gy, gz = np.zeros(n), np.zeros(n), np.zeros(n)
    _usg_lib().orc_usg_global_xyz(C.byref(OP.s), n, _ptr(_i(cell), _ip),
                                  _ptr(_d(lx), _dp), _ptr(_d(ly), _dp), _ptr(_d(lz), _dp),
                                  _ptr(gx, _dp), _ptr(gy, _dp), _ptr(gz, _dp))
    return gx, gy, gz
