"""Host-side mirror of MODPATH-RW's tracking engine over the librwpt C ABI.

``Engine`` plays the role of ParticleTrackingEngineType
(ParticleTrackingEngine.f90:393-492 Initialize, pr_RWPTrackPath :1051-1254)
driven the way MPathRW.F90:1119-1725 drives it: set the flow and the
dispersion once, load the particle groups, then ``advance(t_stop)`` once per
tracking interval.  All compute runs in librwpt.so (HIP, gfx950); this module
only moves numpy arrays across the C ABI and raises on error codes.  There is
no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_LIBPATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "librwpt.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)


class RwptGrid(C.Structure):
    _fields_ = [("ncol", C.c_int32), ("nrow", C.c_int32), ("nlay", C.c_int32),
                ("delr", _dp), ("delc", _dp), ("top", _dp), ("bot", _dp)]


class RwptUsgGrid(C.Structure):
    _fields_ = [("ncell", C.c_int32), ("nlay", C.c_int32), ("nja", C.c_int64),
                ("ia", _i32p), ("ja", _i32p), ("layer", _i32p), ("cell_x", _dp),
                ("cell_y", _dp), ("del_x", _dp), ("del_y", _dp), ("top", _dp),
                ("bot", _dp), ("cell_type", _i32p)]


class RwptOptions(C.Structure):
    _fields_ = [("advection_kind", C.c_int32), ("time_step_kind", C.c_int32),
                ("time_step_params", C.c_double * 2),
                ("dimension_mask", C.c_int32 * 3),
                ("backward_tracking", C.c_int32), ("stop_zone", C.c_int32),
                ("stop_at_weak_sinks", C.c_int32),
                ("stop_at_weak_sources", C.c_int32),
                ("steady_state", C.c_int32), ("precision", C.c_int32),
                ("base_seed", C.c_uint64 * 2)]


class RwptKdeOptions(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("bin_size", C.c_double * 3),
                ("nbins", C.c_int32 * 3), ("mode", C.c_int32),
                ("weighted", C.c_int32), ("scaling_factor", C.c_double),
                ("unit_volume", C.c_int32), ("n_opt_loops", C.c_int32),
                ("isotropic", C.c_int32), ("skip_error_convergence", C.c_int32),
                ("rel_error_convergence", C.c_double),
                ("initial_smoothing", C.c_int32),
                ("initial_smoothing_factor", C.c_double),
                ("bound_kernel_size", C.c_int32), ("min_h_over_delta", C.c_double),
                ("max_h_over_delta", C.c_double),
                ("effective_weight_format", C.c_int32),
                ("kernel_database", C.c_int32), ("delta_h_over_delta", C.c_double),
                ("member", C.c_int32), ("initial", C.c_int32)]


class RwptAdjointOpts(C.Structure):
    _fields_ = [("graze_cut", C.c_double), ("reverse_windows", C.c_int32),
                ("tangent_windows", C.c_int32), ("frozen_noise", C.c_int32),
                ("diagnostics", C.c_int32)]


EXPORTS = {
    "rwpt_default_options": ([C.POINTER(RwptOptions)], None),
    "rwpt_create": ([C.POINTER(RwptGrid), C.POINTER(RwptOptions),
                     C.POINTER(C.c_void_p)], C.c_int),
    "rwpt_create_ensemble": ([C.POINTER(RwptGrid), C.POINTER(RwptOptions), C.c_int32,
                              C.POINTER(C.c_void_p)], C.c_int),
    "rwpt_create_usg": ([C.POINTER(RwptUsgGrid), C.POINTER(RwptOptions),
                         C.POINTER(C.c_void_p)], C.c_int),
    "rwpt_set_flow_usg": ([C.c_void_p, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p, C.c_int32],
                          C.c_int),
    "rwpt_usg_unit_count": ([C.c_void_p], C.c_int64),
    "rwpt_get_usg_tables": ([C.c_void_p, _dp], C.c_int),
    "rwpt_destroy": ([C.c_void_p], C.c_int),
    "rwpt_last_error": ([C.c_void_p], C.c_char_p),
    "rwpt_set_flow": ([C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p,
                       C.c_int32], C.c_int),
    "rwpt_set_cell_extras": ([C.c_void_p, _dp, _i32p, _i32p, _dp, _dp], C.c_int),
    "rwpt_set_saturated_top": ([C.c_void_p, _dp], C.c_int),
    "rwpt_set_graph": ([C.c_void_p, C.c_int32, C.c_int64], C.c_int),
    "rwpt_set_dispersion": ([C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _dp,
                             _dp, _dp], C.c_int),
    "rwpt_load_particles": ([C.c_void_p, C.c_int64, _i32p, _dp, _dp, _dp, _dp,
                             _i32p, _u64p, C.c_int64], C.c_int),
    "rwpt_get_particles": ([C.c_void_p, _i32p, _dp, _dp, _dp, _dp, _i32p,
                            _u64p], C.c_int),
    "rwpt_particle_count": ([C.c_void_p], C.c_int64),
    "rwpt_advance": ([C.c_void_p, C.c_double, _i64p, _i64p], C.c_int),
    "rwpt_set_time": ([C.c_void_p, C.c_double], C.c_int),
    "rwpt_set_steady_state": ([C.c_void_p, C.c_int32], C.c_int),
    "rwpt_set_tuning": ([C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32],
                        C.c_int),
    "rwpt_last_launches": ([C.c_void_p], C.c_int),
    "rwpt_get_substeps": ([C.c_void_p, _i64p], C.c_int),
    "rwpt_autotune": ([C.c_void_p, C.c_double, _i32p, _i32p, _i32p, _i32p, _dp,
                       _i64p], C.c_int),
    "rwpt_global_xyz": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "rwpt_kde": ([C.c_void_p, C.POINTER(RwptKdeOptions), _dp, _dp, _dp], C.c_int),
    "rwpt_kde_device": ([C.c_void_p, C.POINTER(RwptKdeOptions), C.c_void_p, C.c_void_p],
                        C.c_int),
    "rwpt_kde_info": ([C.c_void_p, _dp, _dp, _dp, _dp, _i32p], C.c_int),
    "rwpt_write_timeseries_header": ([C.c_char_p, C.c_int32, C.c_double, C.c_double,
                                      C.c_double, C.c_double], C.c_int),
    "rwpt_write_timeseries": ([C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_double,
                               _i32p, _i32p, _i32p, C.POINTER(C.c_uint8)], C.c_int),
    "rwpt_write_endpoints": ([C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_double,
                              C.c_double, C.c_double, C.c_double, C.c_int32,
                              C.POINTER(C.c_char_p), _i32p, _i32p, _i32p, _i32p, _i32p,
                              _dp, _dp, _dp, _dp, _i32p, _dp], C.c_int),
    "rwpt_write_observation": ([C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                C.c_int32, _i32p, C.c_int32, C.c_int32, _i32p, _i32p,
                                _i32p, _dp, C.c_double, C.POINTER(C.c_uint8)], C.c_int),
    "rwpt_kde_write": ([C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                        C.c_double, C.c_int32, C.POINTER(RwptKdeOptions), _dp, _dp,
                        C.c_double], C.c_int),
    "rwpt_get_cell_tables": ([C.c_void_p, _dp], C.c_int),
    "rwpt_adjoint": ([C.c_void_p, C.c_int32, _dp, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp,
                      _dp, _dp, _dp, _dp, _dp], C.c_int),
    "rwpt_adjoint_options": ([C.c_void_p, C.POINTER(RwptAdjointOpts)], C.c_int),
    "rwpt_adjoint_stats": ([C.c_void_p, _dp], C.c_int),
    "rwpt_kernel_time": ([C.c_void_p, C.c_int32, _dp, _i64p], C.c_int),
    "rwpt_reset_kernel_time": ([C.c_void_p], C.c_int),
}


class RwptError(RuntimeError):
    pass


def library_path() -> str:
    return _LIBPATH


def lib():
    """Loads librwpt.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIBPATH):
            raise RwptError(f"{_LIBPATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(_LIBPATH)
        for name, (args, res) in EXPORTS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def default_options(**kw) -> RwptOptions:
    o = RwptOptions()
    lib().rwpt_default_options(C.byref(o))
    for k, v in kw.items():
        if k in ("time_step_params",):
            o.time_step_params[0], o.time_step_params[1] = v
        elif k == "dimension_mask":
            for d in range(3):
                o.dimension_mask[d] = v[d]
        elif k == "base_seed":
            o.base_seed[0], o.base_seed[1] = v
        else:
            setattr(o, k, v)
    return o


def write_timeseries_header(path, tracking_direction, reference_time, origin, rotation):
    """Replaces path with MODPATH-RW's timeseries header."""
    rc = lib().rwpt_write_timeseries_header(os.fsencode(path), int(tracking_direction),
                                            float(reference_time), float(origin[0]),
                                            float(origin[1]), float(rotation))
    if rc != 0:
        raise RwptError(f"rwpt_write_timeseries_header failed ({rc})")


def kde_write(path, origin, bin_size, nbins, density, histogram, time_id=1, time=0.0,
              group_id=1, fmt=0, column_format=0, histogram_scale=1.0, append=False):
    """Writes the density rows of a reconstruction as MODPATH-RW's gpkde
    output file holds them (rwpt_kde_write; no engine or device needed)."""
    o = RwptKdeOptions()
    for d in range(3):
        o.origin[d] = origin[d]
        o.bin_size[d] = bin_size[d]
        o.nbins[d] = nbins[d]
    dens = _d(density)
    hist = _d(histogram)
    if dens.size != int(np.prod(nbins)) or hist.size != dens.size:
        raise RwptError("density and histogram must hold nbins cells")
    rc = lib().rwpt_kde_write(os.fsencode(path), int(bool(append)), fmt, column_format,
                              time_id, time, group_id, C.byref(o), _p(dens, _dp),
                              _p(hist, _dp), histogram_scale)
    if rc != 0:
        raise RwptError(f"rwpt_kde_write failed ({rc})")


class Engine:
    """One particle-tracking engine resident on the current HIP device."""

    def __init__(self, problem, precision: int = 0, **opt):
        """problem: a Problem, an ensemble from problems.stack (its members
        share the first member's grid), or a UsgProblem (an unstructured
        grid of rectangular cells)."""
        P = problem
        self.problem = P
        self.members = getattr(P, "members", 1)
        self.unstructured = hasattr(P, "ja")
        self._keep = []
        if self.unstructured:
            g = RwptUsgGrid()
            g.ncell, g.nlay, g.nja = int(P.ncell), int(P.nlay), int(np.asarray(P.ja).size)
            ints = [_i(P.ia), _i(P.ja), _i(P.layer), _i(P.layer_type)]
            dbls = [_d(P.cell_x), _d(P.cell_y), _d(P.del_x), _d(P.del_y), _d(P.top),
                    _d(P.bot)]
            self._keep += ints + dbls
            g.ia, g.ja, g.layer, g.cell_type = [_p(a, _i32p) for a in ints]
            g.cell_x, g.cell_y, g.del_x, g.del_y, g.top, g.bot = [_p(a, _dp) for a in dbls]
            ncm = int(P.ncell)
        else:
            g = RwptGrid()
            g.ncol, g.nrow, g.nlay = P.ncol, P.nrow, P.nlay
            ncm = P.ncol * P.nrow * P.nlay
            arrs = [_d(P.delr), _d(P.delc), _d(P.top)[:ncm].copy(), _d(P.bot)[:ncm].copy()]
            self._keep += arrs
            g.delr, g.delc, g.top, g.bot = [_p(a, _dp) for a in arrs]
        stops = {k: int(getattr(P, a, d)) for k, a, d in (
            ("stop_zone", "stop_zone", 0), ("stop_at_weak_sinks", "stop_weak_sinks", 0),
            ("stop_at_weak_sources", "stop_weak_sources", 0),
            ("steady_state", "steady_state", 1))}
        stops.update(opt)
        o = default_options(advection_kind=P.advection_kind,
                            time_step_kind=P.time_step_kind,
                            time_step_params=P.ts_params,
                            dimension_mask=P.dim_mask,
                            backward_tracking=P.backward, precision=precision,
                            **stops)
        h = C.c_void_p()
        if self.unstructured:
            rc = lib().rwpt_create_usg(C.byref(g), C.byref(o), C.byref(h))
            what = "rwpt_create_usg"
        else:
            rc = lib().rwpt_create_ensemble(C.byref(g), C.byref(o), int(self.members),
                                            C.byref(h))
            what = "rwpt_create_ensemble"
        if rc != 0:
            raise RwptError(f"{what} failed ({rc})")
        self.h = h
        self.n = 0
        self.ncell_total = ncm * self.members
        self.set_flow(P)
        self.set_dispersion(P)
        if not self.unstructured and getattr(P, "has_extras", False):
            self.set_cell_extras(P.head, P.layer_type, P.zone, P.source_flow, P.sink_flow)

    # -- error handling -------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = lib().rwpt_last_error(self.h).decode()
            raise RwptError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            lib().rwpt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs ---------------------------------------------------------------
    def set_flow(self, P):
        if self.unstructured:
            a = [_d(P.flowja), _d(P.bflow), _d(P.storage), _d(P.porosity),
                 _d(P.retardation)]
            ib, ic = _i(P.ibound), _i(P.icbound)
            rc = lib().rwpt_set_flow_usg(self.h, *[_p(x, _dp) for x in a], _p(ib, _i32p),
                                         _p(ic, _i32p), int(P.default_icbound))
            self._check(rc, "rwpt_set_flow_usg")
            # the problem brings its heads, zones and source and sink flows
            self.set_cell_extras(P.head, None, P.zone, P.source, P.sink)
        else:
            a = [_d(P.frf), _d(P.fff), _d(P.flf), _d(P.bflow), _d(P.porosity),
                 _d(P.retardation)]
            ib, ic = _i(P.ibound), _i(P.icbound)
            rc = lib().rwpt_set_flow(self.h, *[_p(x, _dp) for x in a], _p(ib, _i32p),
                                     _p(ic, _i32p), int(P.default_icbound))
            self._check(rc, "rwpt_set_flow")
        self.set_saturated_top(getattr(P, "saturated_top", None))

    def set_graph(self, launches: int, max_list: int = 0):
        """rwpt_set_graph: continue each interval's walk, once at most
        max_list particles are unfinished (0: from the start), as replays of
        a captured graph of `launches` walk launches (0: eager launches)."""
        self._check(lib().rwpt_set_graph(self.h, int(launches), int(max_list)),
                    "rwpt_set_graph")

    def set_saturated_top(self, saturated_top=None):
        """rwpt_set_saturated_top: Grid%SaturatedTop of the time step (one
        value per cell), the top of the global z; None: the cell tops."""
        a = _d(saturated_top)
        self._check(lib().rwpt_set_saturated_top(self.h, _p(a, _dp)),
                    "rwpt_set_saturated_top")

    def set_dispersion(self, P):
        a = [_d(P.aL), _d(P.aT), _d(P.aLV), _d(P.aTV), _d(P.dmeff)]
        rc = lib().rwpt_set_dispersion(self.h, int(P.disp_model), int(P.disp_uniform),
                                       *[_p(x, _dp) for x in a])
        self._check(rc, "rwpt_set_dispersion")

    def set_cell_extras(self, head=None, layer_type=None, zone=None, source_flow=None,
                        sink_flow=None):
        """rwpt_set_cell_extras: heads and layer types of convertible layers,
        zones, weak source and sink flows (each None or one value per cell)."""
        a = [_d(head), _i(layer_type), _i(zone), _d(source_flow), _d(sink_flow)]
        rc = lib().rwpt_set_cell_extras(self.h, _p(a[0], _dp), _p(a[1], _i32p),
                                        _p(a[2], _i32p), _p(a[3], _dp), _p(a[4], _dp))
        self._check(rc, "rwpt_set_cell_extras")
        self._keep_extras = a

    def unit_count(self) -> int:
        """Tracking units of an unstructured engine (sub-cells of refined
        cells count one each)."""
        return int(lib().rwpt_usg_unit_count(self.h))

    def usg_tables(self) -> np.ndarray:
        """rwpt_get_usg_tables: (units, 126) checking copy of the unit
        records."""
        out = np.zeros((self.unit_count(), 126))
        self._check(lib().rwpt_get_usg_tables(self.h, _p(out, _dp)), "rwpt_get_usg_tables")
        return out

    def load_particles(self, cell, lx, ly, lz, t=None, status=None, rng=None,
                       stream_offset=0):
        n = int(np.asarray(cell).size)
        cell, lx, ly, lz = _i(cell), _d(lx), _d(ly), _d(lz)
        t = _d(np.zeros(n) if t is None else t)
        status = _i(np.ones(n, np.int32) if status is None else status)
        rng = None if rng is None else np.ascontiguousarray(rng, dtype=np.uint64)
        rc = lib().rwpt_load_particles(self.h, n, _p(cell, _i32p), _p(lx, _dp),
                                       _p(ly, _dp), _p(lz, _dp), _p(t, _dp),
                                       _p(status, _i32p), _p(rng, _u64p),
                                       int(stream_offset))
        self._check(rc, "rwpt_load_particles")
        self.n = n

    # -- stepping -------------------------------------------------------------
    def set_time(self, t):
        self._check(lib().rwpt_set_time(self.h, float(t)), "rwpt_set_time")

    def set_steady_state(self, steady):
        """rwpt_set_steady_state: FlowModelData%SteadyState of the loaded
        time step."""
        self._check(lib().rwpt_set_steady_state(self.h, int(bool(steady))),
                    "rwpt_set_steady_state")

    def set_tuning(self, budget=64, block=256, sort=True, min_waves=1):
        self._check(lib().rwpt_set_tuning(self.h, int(budget), int(block), int(sort),
                                          int(min_waves)), "rwpt_set_tuning")

    def autotune(self, t_probe):
        """rwpt_autotune: keeps the fastest launch shape for advances like
        the one to t_probe; the particle state is left as it was."""
        v = [C.c_int32(0) for _ in range(4)]
        ms = C.c_double(0)
        steps = C.c_int64(0)
        self._check(lib().rwpt_autotune(self.h, float(t_probe), *[C.byref(x) for x in v],
                                        C.byref(ms), C.byref(steps)), "rwpt_autotune")
        return {"budget": v[0].value, "block": v[1].value, "sort": v[2].value,
                "min_waves": v[3].value, "ms": ms.value, "trial_substeps": steps.value}

    def adjoint(self, t_windows, obs_xyz, target, bandwidth, mass=None, gradient=True):
        """rwpt_adjoint: advances through the window times, returns the misfit
        sum_o (c_o - target_o)^2 of the Gaussian-smoothed concentrations at
        the observation points and, with gradient, its pathwise derivative
        with respect to every cell's aLH, aTH, Dm (arrays of ncell) and to a
        common porosity scale."""
        tw = np.ascontiguousarray(t_windows, dtype=np.float64)
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        tgt = np.ascontiguousarray(target, dtype=np.float64)
        if tgt.size != obs.shape[0]:
            raise ValueError("one target per observation point")
        bw = np.ascontiguousarray(bandwidth, dtype=np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        conc = np.zeros(obs.shape[0])
        J, ms, gs = C.c_double(0), C.c_double(0), C.c_double(0)
        nc = self.ncell_total
        g = [np.zeros(nc) for _ in range(3)] if gradient else [None, None, None]
        rc = lib().rwpt_adjoint(self.h, int(tw.size), _p(tw, _dp), int(obs.shape[0]),
                                _p(obs, _dp), _p(tgt, _dp), _p(bw, _dp), _p(m, _dp),
                                C.byref(J), _p(conc, _dp), _p(g[0], _dp), _p(g[1], _dp),
                                _p(g[2], _dp), C.byref(gs) if gradient else None,
                                C.byref(ms))
        self._check(rc, "rwpt_adjoint")
        out = {"misfit": J.value, "conc": conc, "ms": ms.value}
        if gradient:
            out.update({"grad_al": g[0], "grad_at": g[1], "grad_dm": g[2],
                        "grad_porosity_scale": gs.value})
        return out

    def adjoint_options(self, graze_cut=0.0, reverse_windows=0, tangent_windows=0,
                        frozen_noise=False, diagnostics=False):
        """rwpt_adjoint_options: exits with kappa < graze_cut lose their
        exit-time derivative; reverse_windows > 0 sweeps only the last
        windows; tangent_windows > 0 restarts the adjoint from the final
        time's every that many windows (straight-through); frozen_noise
        drops the random displacements' derivative with respect to the
        sub-step duration; diagnostics tallies the reverse sweep."""
        o = RwptAdjointOpts(float(graze_cut), int(reverse_windows), int(tangent_windows),
                            int(bool(frozen_noise)), int(bool(diagnostics)))
        self._check(lib().rwpt_adjoint_options(self.h, C.byref(o)), "rwpt_adjoint_options")

    ADJ_KINDS = ("substep", "clipped", "exit", "grazing_exit", "rare")
    ADJ_KAPPA_BINS = ("<1e-10", "1e-10..1e-8", "1e-8..1e-6", "1e-6..1e-4", "1e-4..1e-2",
                      "1e-2..1e-1", ">=1e-1")

    def adjoint_stats(self):
        """rwpt_adjoint_stats of the last gradient taken with diagnostics on,
        by iteration kind (substep, clipped, exit, grazing_exit with
        kappa < 1e-3, rare): iterations, sum and maximum of log10 adjoint
        growth, iterations growing by more than 10; exits by kappa decade;
        whole-horizon growth of the particles."""
        v = np.zeros(33)
        self._check(lib().rwpt_adjoint_stats(self.h, _p(v, _dp)), "rwpt_adjoint_stats")
        kinds = {k: {"iterations": int(v[i]), "log10_growth_sum": float(v[5 + i]),
                     "log10_growth_max": float(v[10 + i]), "growth_over_10": int(v[15 + i])}
                 for i, k in enumerate(self.ADJ_KINDS)}
        return {"kinds": kinds,
                "exit_kappa": dict(zip(self.ADJ_KAPPA_BINS, (int(x) for x in v[20:27]))),
                "exits_cut": int(v[27]), "particles": int(v[28]),
                "particle_log10_growth_max": float(v[29]),
                "particles_growth_over": {"1e3": int(v[30]), "1e6": int(v[31]),
                                          "1e9": int(v[32])}}

    def write_observation(self, path, style, cells, time_point, time_step, ids, group,
                          group_solute, mass=None, flow_rate=0.0, fmt=0, append=True,
                          select=None):
        """Observation records of the particles in `cells` (rwpt_write_observation):
        style 1 resident (active particles), 2 sink (stopped at a weak sink)."""
        cl, idv, grp, sol = _i(cells), _i(ids), _i(group), _i(group_solute)
        m = None if mass is None else _d(mass)
        sel = None if select is None else np.ascontiguousarray(select, dtype=np.uint8)
        rc = lib().rwpt_write_observation(
            self.h, os.fsencode(path), int(bool(append)), int(style), int(fmt), cl.size,
            _p(cl, _i32p), int(time_point), int(time_step), _p(idv, _i32p), _p(grp, _i32p),
            _p(sol, _i32p), _p(m, _dp), float(flow_rate),
            None if sel is None else sel.ctypes.data_as(C.POINTER(C.c_uint8)))
        self._check(rc, "rwpt_write_observation")

    def write_timeseries(self, path, time_point, time_step, time, sequence, ids, group,
                         select=None):
        """Appends the timeseries records of the particles (rwpt_write_timeseries)."""
        seq, idv, grp = _i(sequence), _i(ids), _i(group)
        sel = None if select is None else np.ascontiguousarray(select, dtype=np.uint8)
        rc = lib().rwpt_write_timeseries(
            self.h, os.fsencode(path), int(time_point), int(time_step), float(time),
            _p(seq, _i32p), _p(idv, _i32p), _p(grp, _i32p),
            None if sel is None else sel.ctypes.data_as(C.POINTER(C.c_uint8)))
        self._check(rc, "rwpt_write_timeseries")

    def write_endpoints(self, path, fmt, tracking_direction, reference_time, origin,
                        rotation, group_names, group_solute, group, sequence, ids,
                        initial_cell, initial_lx, initial_ly, initial_lz, initial_time,
                        initial_face=None, mass=None):
        """Writes the endpoint file (rwpt_write_endpoints)."""
        names = (C.c_char_p * len(group_names))(*[n.encode() for n in group_names])
        arr_i = [_i(group_solute), _i(group), _i(sequence), _i(ids), _i(initial_cell)]
        arr_d = [_d(initial_lx), _d(initial_ly), _d(initial_lz), _d(initial_time)]
        face = None if initial_face is None else _i(initial_face)
        m = None if mass is None else _d(mass)
        rc = lib().rwpt_write_endpoints(
            self.h, os.fsencode(path), int(fmt), int(tracking_direction),
            float(reference_time), float(origin[0]), float(origin[1]), float(rotation),
            len(group_names), names, *[_p(a, _i32p) for a in arr_i],
            *[_p(a, _dp) for a in arr_d], _p(face, _i32p), _p(m, _dp))
        self._check(rc, "rwpt_write_endpoints")

    def last_launches(self):
        return lib().rwpt_last_launches(self.h)

    def advance(self, t_stop):
        act = C.c_int64(0)
        ns = C.c_int64(0)
        rc = lib().rwpt_advance(self.h, float(t_stop), C.byref(act), C.byref(ns))
        self._check(rc, "rwpt_advance")
        return act.value, ns.value

    # -- outputs --------------------------------------------------------------
    def particles(self):
        n = self.n
        out = {"cell": np.zeros(n, np.int32), "lx": np.zeros(n), "ly": np.zeros(n),
               "lz": np.zeros(n), "t": np.zeros(n), "status": np.zeros(n, np.int32),
               "rng": np.zeros(2 * n, np.uint64)}
        rc = lib().rwpt_get_particles(self.h, _p(out["cell"], _i32p),
                                      _p(out["lx"], _dp), _p(out["ly"], _dp),
                                      _p(out["lz"], _dp), _p(out["t"], _dp),
                                      _p(out["status"], _i32p), _p(out["rng"], _u64p))
        self._check(rc, "rwpt_get_particles")
        sub = np.zeros(n, np.int64)
        self._check(lib().rwpt_get_substeps(self.h, _p(sub, _i64p)), "rwpt_get_substeps")
        out["substeps"] = sub
        return out

    def global_xyz(self):
        n = self.n
        gx, gy, gz = np.zeros(n), np.zeros(n), np.zeros(n)
        rc = lib().rwpt_global_xyz(self.h, _p(gx, _dp), _p(gy, _dp), _p(gz, _dp))
        self._check(rc, "rwpt_global_xyz")
        return gx, gy, gz

    def cell_tables(self):
        out = np.zeros((self.problem.ncell, 104))
        self._check(lib().rwpt_get_cell_tables(self.h, _p(out, _dp)),
                    "rwpt_get_cell_tables")
        return out

    def kernel_time(self, which=0):
        ms = C.c_double(0)
        n = C.c_int64(0)
        self._check(lib().rwpt_kernel_time(self.h, which, C.byref(ms), C.byref(n)),
                    "rwpt_kernel_time")
        return ms.value, n.value

    def reset_kernel_time(self):
        self._check(lib().rwpt_reset_kernel_time(self.h), "rwpt_reset_kernel_time")

    @staticmethod
    def kde_options(origin, bin_size, nbins, mode=0, weighted=False, scaling_factor=1.0,
                    unit_volume=False, n_opt_loops=0, isotropic=False, skip_error=False,
                    rel_error=0.02, initial_smoothing=0, smoothing_factor=1.0,
                    bound_kernel_size=0, min_h_over_delta=0.0, max_h_over_delta=0.0,
                    effective_weight_format=0, kernel_database=False,
                    delta_h_over_delta=0.0, member=-1, initial=False):
        """rwpt_kde_options from the options of the gpkde input file
        (ReadGPKDEData, ModpathSimulationData.f90:816-1273)."""
        o = RwptKdeOptions()
        for d in range(3):
            o.origin[d] = origin[d]
            o.bin_size[d] = bin_size[d]
            o.nbins[d] = nbins[d]
        o.mode = mode
        o.weighted = int(bool(weighted))
        o.scaling_factor = scaling_factor
        o.unit_volume = int(bool(unit_volume))
        o.n_opt_loops = n_opt_loops
        o.isotropic = int(bool(isotropic))
        o.skip_error_convergence = int(bool(skip_error))
        o.rel_error_convergence = rel_error
        o.initial_smoothing = initial_smoothing
        o.initial_smoothing_factor = smoothing_factor
        o.bound_kernel_size = bound_kernel_size
        o.min_h_over_delta = min_h_over_delta
        o.max_h_over_delta = max_h_over_delta
        o.effective_weight_format = effective_weight_format
        o.kernel_database = int(bool(kernel_database))
        o.delta_h_over_delta = delta_h_over_delta
        o.member = member
        o.initial = int(bool(initial))
        return o

    def kde(self, origin, bin_size, nbins, mode=0, mass=None, **kw):
        """gpkde reconstruction of the active particles; the keywords are
        those of kde_options.  Returns (density, histogram), x fastest."""
        o = self.kde_options(origin, bin_size, nbins, mode=mode, weighted=mass is not None,
                             **kw)
        nb = int(np.prod(nbins))
        dens = np.zeros(nb)
        hist = np.zeros(nb)
        m = _d(mass)
        rc = lib().rwpt_kde(self.h, C.byref(o), _p(m, _dp), _p(dens, _dp), _p(hist, _dp))
        self._check(rc, "rwpt_kde")
        return dens, hist

    def kde_device(self, density_dev, histogram_dev, origin, bin_size, nbins, mode=1, **kw):
        """rwpt_kde_device: the one-pass reconstruction into device buffers
        (addresses as integers), allocating nothing."""
        o = self.kde_options(origin, bin_size, nbins, mode=mode, **kw)
        rc = lib().rwpt_kde_device(self.h, C.byref(o), C.c_void_p(density_dev),
                                   C.c_void_p(histogram_dev))
        self._check(rc, "rwpt_kde_device")

    def kde_info(self):
        v = [C.c_double(0) for _ in range(4)]
        loops = C.c_int32(0)
        self._check(lib().rwpt_kde_info(self.h, *[C.byref(x) for x in v], C.byref(loops)),
                    "rwpt_kde_info")
        return {"bandwidth": v[0].value, "points": v[1].value, "npoints": v[2].value,
                "eff_mass": v[3].value, "loops": loops.value}
