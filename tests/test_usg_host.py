"""Host-side checks of rwpt_create_usg (no device needed: every rejection
happens before the first HIP call)."""
import ctypes as C

import numpy as np

from rwpt import engine
from usg_synth import Quadtree


def _grid(Q, keep, ja=None):
    A = Q.cell_arrays()
    g = engine.RwptUsgGrid()
    g.ncell, g.nlay = Q.ncell, Q.nlay
    ia = np.ascontiguousarray(Q.ia, np.int32)
    ja = np.ascontiguousarray(Q.ja if ja is None else ja, np.int32)
    g.nja = ja.size
    arrs = [ia, ja, np.ascontiguousarray(A["layer"], np.int32)]
    dbl = [np.ascontiguousarray(A[k], np.float64)
           for k in ("cell_x", "cell_y", "del_x", "del_y", "top", "bot")]
    keep += arrs + dbl
    g.ia, g.ja, g.layer = [a.ctypes.data_as(C.POINTER(C.c_int32)) for a in arrs]
    g.cell_x, g.cell_y, g.del_x, g.del_y, g.top, g.bot = [
        a.ctypes.data_as(C.POINTER(C.c_double)) for a in dbl]
    return g


def test_create_usg_rejects_bad_input_on_the_host():
    lib = engine.lib()
    Q = Quadtree(3, 4, {(1, 1)}, nlay=1)
    keep = []
    h = C.c_void_p()
    o = engine.default_options()
    # rows must be led by their own cell
    bad = Q.ja.copy()
    bad[0] = 2
    assert lib.rwpt_create_usg(C.byref(_grid(Q, keep, bad)), C.byref(o), C.byref(h)) == -1
    # the unstructured walk is FP64 only
    o1 = engine.default_options(precision=1)
    assert lib.rwpt_create_usg(C.byref(_grid(Q, keep)), C.byref(o1), C.byref(h)) == -5
    assert lib.rwpt_create_usg(None, C.byref(o), C.byref(h)) == -1
    assert not h.value


def test_create_usg_rejects_spacing_of_no_quadtree():
    """A cell whose face meets a cell of an eighth of its height is no
    smoothed quadtree (ComputeFaceAssignments stops there)."""
    lib = engine.lib()
    Q = Quadtree(2, 2, {(0, 1)}, nlay=1)
    # the upper-left child of the refined cell, east of cell 1, made 1/8 high
    Q.del_y = Q.del_y.copy()
    Q.del_y[1] = 0.125
    keep = []
    h = C.c_void_p()
    o = engine.default_options()
    assert lib.rwpt_create_usg(C.byref(_grid(Q, keep)), C.byref(o), C.byref(h)) == -5
    assert not h.value
