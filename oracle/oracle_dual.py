"""ctypes binding of the dual-number oracle (liborc_dual.so) -- TEST
INFRASTRUCTURE.

The C restatement of rwpt_oracle.c compiled with every double a dual number
(rwpt_oracle_dual.cpp): forward-mode tangents of the particles' positions
and times in ND parameter directions, given as tangents of the problem's
arrays.  The FP64 reference gradient of the adjoint tests: the misfit's
directional derivative along a direction equals the dot product of the
device's reverse-mode gradient with that direction.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import oracle as O

ND = 4                # tangent directions of the build (ORC_ND)
W = 1 + ND            # doubles per dual value
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)


class OrcDual(C.Structure):
    _fields_ = [("v", C.c_double), ("d", C.c_double * ND)]


def _fields():
    out = []
    for name, t in O.OrcProblem._fields_:
        out.append((name, C.c_double * (2 * W)) if name == "ts_params" else (name, t))
    return out


class OrcDualProblem(C.Structure):
    _fields_ = _fields()


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc_dual.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc_dual.so is not built (run build())")
        L = C.CDLL(path)
        L.orcd_advance.argtypes = [C.POINTER(OrcDualProblem), C.c_long, _ip, _dp, _dp,
                                   _dp, _dp, _ip, _u64p, OrcDual, OrcDual,
                                   C.POINTER(C.c_longlong), C.c_int]
        L.orcd_advance.restype = C.c_int
        L.orcd_global_xyz.argtypes = [C.POINTER(OrcDualProblem), C.c_long, _ip, _dp, _dp,
                                      _dp, _dp, _dp, _dp]
        _LIB = L
    return _LIB


def dual(values, tangents=None):
    """(n, 1 + ND) array: values and tangents (n, ND) or zero."""
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    out = np.zeros((v.size, W))
    out[:, 0] = v
    if tangents is not None:
        out[:, 1:] = np.asarray(tangents, dtype=np.float64).reshape(v.size, ND)
    return out


class DualProblem:
    """The oracle's problem with tangents: `tangents` maps array names of the
    problem (aL, aT, dmeff, porosity, ...) to (n, ND) tangent arrays."""

    def __init__(self, P, tangents):
        base = O.OracleProblem(P)
        self._keep = {}
        s = OrcDualProblem()
        for name, t in O.OrcProblem._fields_:
            if name == "ts_params":
                s.ts_params[0] = base.s.ts_params[0]
                s.ts_params[W] = base.s.ts_params[1]
            elif t is O._dp:
                a = base._keep.get(name)
                if a is None:
                    setattr(s, name, None)
                    continue
                d = dual(a, tangents.get(name))
                self._keep[name] = d
                setattr(s, name, d.ctypes.data_as(_dp))
            else:
                setattr(s, name, getattr(base.s, name))
        self._base = base
        self.s = s


def _pd(a):
    return a.ctypes.data_as(_dp)


def advance(DP, state, t_engine, t_stop):
    """state: cell, status, rng, substeps as for oracle.advance and lx, ly,
    lz, t as (n, 1 + ND) dual arrays, modified in place."""
    n = state["cell"].size
    te, ts = OrcDual(float(t_engine)), OrcDual(float(t_stop))
    return lib().orcd_advance(C.byref(DP.s), n, state["cell"].ctypes.data_as(_ip),
                              _pd(state["lx"]), _pd(state["ly"]), _pd(state["lz"]),
                              _pd(state["t"]), state["status"].ctypes.data_as(_ip),
                              state["rng"].ctypes.data_as(_u64p), te, ts,
                              state["substeps"].ctypes.data_as(C.POINTER(C.c_longlong)), 1)


def new_state(cell, lx, ly, lz):
    st = O.new_state(cell, lx, ly, lz)
    for k in ("lx", "ly", "lz", "t"):
        st[k] = dual(st[k])
    return st


def global_xyz(DP, cell, lx, ly, lz):
    """Dual global coordinates (ConvertToModelXYZ) of dual local ones."""
    n = cell.size
    g = [np.zeros((n, W)) for _ in range(3)]
    c = np.ascontiguousarray(cell, dtype=np.int32)
    lib().orcd_global_xyz(C.byref(DP.s), n, c.ctypes.data_as(_ip), _pd(lx), _pd(ly),
                          _pd(lz), _pd(g[0]), _pd(g[1]), _pd(g[2]))
    return g
