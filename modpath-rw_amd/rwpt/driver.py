"""MODPATH-RW's tracking-interval schedule and a driver over the engine.

``schedule`` restates how MPathRW picks the end of every tracking interval
(TIME_STEP_LOOP, MPathRW.F90:850-1733; TRACKING_INTERVAL_LOOP,
:1119-1725): within each MODFLOW time step (end time tsMax, :937-966) the
interval ends at the next timeseries point if it lies within the step, else
at tsMax (:1126-1139); gpkde reconstructions either follow the timeseries
points (gpkdeTimePointOption 0) or have their own times, and then the
earlier of the two ends the interval (:1142-1186).  As in the reference,
when a reconstruction time precedes the pending timeseries point, that
timeseries point's index has already been consumed and is not written
(:1170-1176 reset the flag without restoring nt).  Tracking stops when
the time reaches the stop time (:1730); the Driver also stops after the
first interval that starts with no active or pending particle
(:1716-1722).

``Driver`` runs an Engine over such a schedule, writing what MPathRW writes
and when: the timeseries records and initial-condition reconstruction of
the particles released at time 0 (:1000-1100), the timeseries records,
resident observation records and gpkde density rows at their points, and
the endpoint file after the last interval (:1740-1745).  With a flow source
it loads each MODFLOW time step's flow before the step's first interval, as
TIME_STEP_LOOP calls FlowModelData%LoadTimeStep (MPathRW.F90:850-870).
``prepare_groups`` builds MPathRW's particle arrays from its particle
groups.  All tracking, reconstruction and file formatting runs in librwpt;
this module only orders the calls.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np


@dataclasses.dataclass
class Interval:
    """One pass of TRACKING_INTERVAL_LOOP: advance to `time`."""
    ktime: int                       # MODFLOW time step (1-based)
    time: float                      # maxTime
    time_point: Optional[int]        # timeseries point index nt (1-based) or None
    gpkde_point: Optional[int]       # reconstruction index ngpkde (1-based) or None
    step_end: bool                   # itend: the last interval of the time step


def schedule(step_ends: Sequence[float], stop_time: float,
             time_points: Sequence[float] = (), gpkde_points: Sequence[float] = (),
             gpkde_follows_timeseries: bool = True, stopping_option: int = 2,
             reconstruct: bool = False) -> List[Interval]:
    """The intervals of a forward-tracking run.  step_ends: the end of every
    MODFLOW time step relative to the reference time (TotalTimes -
    ReferenceTime); stopping_option 1 (end of simulation), 2 (extend the
    last step to stop_time) or 3 (stop at stop_time)."""
    out: List[Interval] = []
    nt, ng = 0, 0
    nsteps = len(step_ends)
    for k in range(1, nsteps + 1):
        ts_max = float(step_ends[k - 1])
        if stopping_option == 2:
            if k == nsteps:
                ts_max = stop_time
        elif stopping_option == 3:
            if k == nsteps or ts_max > stop_time:
                ts_max = stop_time
        itend = 0
        time = None
        while itend == 0:
            itend = 1
            max_time = ts_max
            is_ts, is_gp = False, False
            tp, gp = None, None
            if nt + 1 <= len(time_points) and time_points[nt] <= ts_max:
                nt += 1
                max_time = float(time_points[nt - 1])
                itend = 0 if max_time != ts_max else 1
                is_ts, tp = True, nt
            if reconstruct:
                if gpkde_follows_timeseries:
                    if is_ts:
                        ng += 1
                        is_gp, gp = True, ng
                else:
                    mg = 0.0
                    if ng + 1 <= len(gpkde_points) and gpkde_points[ng] <= ts_max:
                        mg = float(gpkde_points[ng])
                    if mg != 0.0:
                        if max_time == min(mg, max_time):
                            if max_time == mg:
                                ng += 1
                                is_gp, gp = True, ng
                        else:
                            # the reconstruction time comes first: the pending
                            # timeseries point is not written
                            is_ts, tp = False, None
                            ng += 1
                            is_gp, gp = True, ng
                            max_time = mg
                            itend = 0 if max_time != ts_max else 1
            out.append(Interval(k, max_time, tp, gp, itend == 1))
            time = max_time
        if time is not None and time >= stop_time:
            break
    return out


@dataclasses.dataclass
class ParticleGroup:
    """A particle group of the simulation file (ModpathSimulationData.f90:
    635-720) with starting locations of input style 1
    (StartingLocationReader.f90:148-240): cells, local coordinates, the
    release-time offsets of the first release and optional particle IDs."""
    name: str
    release_times: Sequence[float]
    cell: np.ndarray
    lx: np.ndarray
    ly: np.ndarray
    lz: np.ndarray
    offset: Optional[np.ndarray] = None
    ids: Optional[np.ndarray] = None
    mass: float = 1.0
    solute: int = 1


def find_face(x, y, z):
    """pr_FindFace (StartingLocationReader.f90:609-635): the face a starting
    location lies on, 0 inside the cell."""
    x, y, z = (np.asarray(a, np.float64) for a in (x, y, z))
    face = np.zeros(x.shape, np.int32)
    for f, on in ((6, z == 1.0), (5, z == 0.0), (4, y == 1.0), (3, y == 0.0),
                  (2, x == 1.0), (1, x == 0.0)):
        face[on] = f          # later assignments win: the reference's order
    return face


def prepare_groups(groups: Sequence[ParticleGroup]) -> dict:
    """The particle arrays of MPathRW's groups in its particle order, as
    ReadAndPrepareLocations builds them (StartingLocationReader.f90:40-145):
    per group the first release (IDs as read, or 1..n; release time = first
    release time + offset), then each later release (IDs continuing from
    the largest so far, release time without the offset); sequence numbers
    run over all groups.  Returns cell, lx, ly, lz, t (release times),
    face, group (1-based), sequence, id, mass and the per-group names and
    solutes."""
    keys = ("cell", "lx", "ly", "lz", "t", "face", "group", "sequence", "id", "mass")
    out = {k: [] for k in keys}
    seq = 0
    for gi, G in enumerate(groups, start=1):
        n = len(G.cell)
        ids = np.arange(1, n + 1) if G.ids is None else np.asarray(G.ids)
        off = np.zeros(n) if G.offset is None else np.asarray(G.offset, np.float64)
        idmax = int(max(0, ids.max())) if n else 0
        for r, t_rel in enumerate(G.release_times):
            if r == 0:
                pid = ids
                t = float(t_rel) + off
            else:
                pid = np.arange(idmax + 1, idmax + n + 1)
                idmax += n
                t = np.full(n, float(t_rel))
            out["cell"].append(np.asarray(G.cell))
            out["lx"].append(np.asarray(G.lx, np.float64))
            out["ly"].append(np.asarray(G.ly, np.float64))
            out["lz"].append(np.asarray(G.lz, np.float64))
            out["t"].append(t)
            out["face"].append(find_face(G.lx, G.ly, G.lz))
            out["group"].append(np.full(n, gi))
            out["sequence"].append(np.arange(seq + 1, seq + n + 1))
            out["id"].append(pid)
            out["mass"].append(np.full(n, float(G.mass)))
            seq += n
    ints = ("cell", "face", "group", "sequence", "id")
    res = {k: (np.concatenate(v) if v else np.zeros(0)).astype(
        np.int32 if k in ints else np.float64) for k, v in out.items()}
    res["names"] = [G.name for G in groups]
    res["solute"] = np.array([G.solute for G in groups], np.int32)
    return res


class Driver:
    """Runs an Engine through a schedule the way MPathRW drives
    ParticleTrackingEngine (a timeseries simulation, type 5), writing the
    outputs at their points."""

    def __init__(self, engine, intervals: Sequence[Interval], *, timeseries_path=None,
                 sequence=None, ids=None, group=None, group_solute=None, mass=None,
                 observations=(), kde_options=None, kde_path=None, kde_format=0,
                 kde_column_format=0, flow_for_step=None, timeseries_option=0,
                 initial_point=True, kde_initial_options=None, endpoint_path=None,
                 endpoint=None):
        """flow_for_step: None (the engine's flow holds for every step) or a
        callable of the 1-based cumulative time step returning a Problem
        (its face flows, boundary flows, porosity, retardation, ibound and
        steady-state flag are set) and the keyword arguments of
        Engine.set_cell_extras (or None).

        timeseries_option: TimeseriesOutputOption, 0 the particles active at
        a timeseries point, 1 every particle, 2 none.  initial_point: the
        records and reconstruction of the particles released at time 0
        before the first interval (MPathRW.F90:1000-1100);
        kde_initial_options: Engine.kde arguments replacing kde_options' for
        that reconstruction (the gpkde file's isotropic initial-condition
        kernels), or False to skip it (gpkdeSkipInitialCondition).
        endpoint_path/endpoint: the endpoint file written after the last
        interval and the keyword arguments of Engine.write_endpoints besides
        the particle arrays (fmt, tracking_direction, reference_time, origin,
        rotation, group_names, initial_face)."""
        self.E = engine
        self.flow_for_step = flow_for_step
        self.intervals = list(intervals)
        self.timeseries_path = timeseries_path
        n = engine.n
        self.sequence = np.arange(1, n + 1, dtype=np.int32) if sequence is None else sequence
        self.ids = self.sequence if ids is None else ids
        self.group = np.ones(n, np.int32) if group is None else group
        self.group_solute = np.ones(1, np.int32) if group_solute is None else group_solute
        # MPathRW reconstructs with the particles' masses (1 unless given)
        self.mass = np.ones(n) if mass is None else mass
        self.observations = list(observations)   # (path, cells) of resident observations
        self.kde_options = kde_options            # dict of Engine.kde arguments
        self.kde_path = kde_path
        self.kde_format = kde_format
        self.kde_column_format = kde_column_format
        self.timeseries_option = timeseries_option
        self.initial_point = initial_point
        self.kde_initial_options = kde_initial_options
        self.endpoint_path = endpoint_path
        self.endpoint = endpoint
        self._kde_written = False

    def _reconstruct(self, time_id, time, options, initial=False):
        from .engine import kde_write
        k = dict(options)
        dens, hist = self.E.kde(mass=self.mass, initial=initial, **k)
        if self.kde_path is not None:
            kde_write(self.kde_path, k["origin"], k["bin_size"], k["nbins"], dens, hist,
                      time_id=time_id, time=time, group_id=1, fmt=self.kde_format,
                      column_format=self.kde_column_format,
                      histogram_scale=k.get("scaling_factor", 1.0),
                      append=self._kde_written)
            self._kde_written = True

    def _load_flow(self, ktime):
        P, extras = self.flow_for_step(ktime)
        self.E.set_flow(P)
        if extras:
            self.E.set_cell_extras(**extras)
        # TrackCell%SteadyState follows the loaded step
        # (ParticleTrackingEngine.f90:893, 1091)
        self.E.set_steady_state(getattr(P, "steady_state", 1))

    def run(self):
        """Advances through every interval; returns the engine's (active,
        sub-steps) after each.  As TRACKING_INTERVAL_LOOP counts the active
        and pending particles of an interval before tracking them, the run
        ends after the first interval that starts with neither
        (MPathRW.F90:1716-1722): one interval after the last particle
        stopped."""
        log = []
        loaded = None
        start = self.E.particles()
        initial = {k: start[k].copy() for k in ("cell", "lx", "ly", "lz", "t")}
        at_zero = (start["status"] == 0) & (start["t"] == 0.0)
        empty = not bool(((start["status"] == 0) | (start["status"] == 1)).any())
        for num, iv in enumerate(self.intervals):
            if self.flow_for_step is not None and iv.ktime != loaded:
                self._load_flow(iv.ktime)
                loaded = iv.ktime
            if num == 0 and self.initial_point:
                # the particles released at time 0, before any tracking
                # (MPathRW.F90:1000-1100)
                if self.timeseries_path is not None and self.timeseries_option != 2:
                    sel = np.ones_like(at_zero) if self.timeseries_option == 1 else at_zero
                    self.E.write_timeseries(self.timeseries_path, 0, iv.ktime, 0.0,
                                            self.sequence, self.ids, self.group, select=sel)
                if (self.kde_options is not None and self.kde_initial_options is not False
                        and at_zero.any()):
                    opts = dict(self.kde_options)
                    opts.update(self.kde_initial_options or {})
                    self._reconstruct(0, 0.0, opts, initial=True)
            active, steps = self.E.advance(iv.time)
            st = self.E.particles()["status"]
            pending = int((st == 0).sum())
            if iv.time_point is not None:
                # a record for every active particle whose tracking in the
                # interval started before the point: one released at the
                # point itself does not move and gets none (pr_FindTimeIndex
                # takes points after the current time,
                # ParticleTrackingEngine.f90:366-390, 1167-1176); option 1
                # writes every particle (MPathRW.F90:1492-1530), resident
                # observation records included
                if self.timeseries_option == 1:
                    sel = np.ones(st.size, bool)
                else:
                    sel = (st == 1) & (initial["t"] < iv.time)
                if self.timeseries_path is not None and self.timeseries_option != 2:
                    self.E.write_timeseries(self.timeseries_path, iv.time_point, iv.ktime,
                                            iv.time, self.sequence, self.ids, self.group,
                                            select=sel)
                for path, cells in self.observations:
                    self.E.write_observation(path, 1, cells, iv.time_point, iv.ktime, self.ids,
                                             self.group, self.group_solute, mass=self.mass,
                                             select=sel)
            if iv.gpkde_point is not None and self.kde_options is not None and active > 0:
                self._reconstruct(iv.gpkde_point, iv.time, self.kde_options)
            log.append((iv, active, steps))
            if empty:
                break
            empty = active == 0 and pending == 0
        if self.endpoint_path is not None:
            ep = dict(self.endpoint or {})
            self.E.write_endpoints(
                self.endpoint_path, ep.pop("fmt", 0), ep.pop("tracking_direction", 1),
                ep.pop("reference_time", 0.0), ep.pop("origin", (0.0, 0.0)),
                ep.pop("rotation", 0.0), ep.pop("group_names"), self.group_solute,
                self.group, self.sequence, self.ids, initial["cell"], initial["lx"],
                initial["ly"], initial["lz"], initial["t"], mass=self.mass, **ep)
        return log
