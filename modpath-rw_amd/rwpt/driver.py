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
(:1170-1176 reset the flag without restoring nt).  Tracking stops after the
interval in which no particle is active or pending (:1716-1722) and when
the time reaches the stop time (:1730).

``Driver`` runs an Engine over such a schedule, writing the timeseries
records, resident observation records and gpkde density rows at the
points the reference writes them.  With a flow source it loads each MODFLOW
time step's flow before the step's first interval, as TIME_STEP_LOOP calls
FlowModelData%LoadTimeStep (MPathRW.F90:850-870).  All tracking, reconstruction and file
formatting runs in librwpt; this module only orders the calls.
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


class Driver:
    """Runs an Engine through a schedule the way MPathRW drives
    ParticleTrackingEngine, writing the outputs at their points."""

    def __init__(self, engine, intervals: Sequence[Interval], *, timeseries_path=None,
                 sequence=None, ids=None, group=None, group_solute=None, mass=None,
                 observations=(), kde_options=None, kde_path=None, kde_format=0,
                 kde_column_format=0, flow_for_step=None):
        """flow_for_step: None (the engine's flow holds for every step) or a
        callable of the 1-based cumulative time step returning a Problem
        (its face flows, boundary flows, porosity, retardation, ibound and
        steady-state flag are set) and the keyword arguments of
        Engine.set_cell_extras (or None)."""
        self.E = engine
        self.flow_for_step = flow_for_step
        self.intervals = list(intervals)
        self.timeseries_path = timeseries_path
        n = engine.n
        self.sequence = np.arange(1, n + 1, dtype=np.int32) if sequence is None else sequence
        self.ids = self.sequence if ids is None else ids
        self.group = np.ones(n, np.int32) if group is None else group
        self.group_solute = np.ones(1, np.int32) if group_solute is None else group_solute
        self.mass = mass
        self.observations = list(observations)   # (path, cells) of resident observations
        self.kde_options = kde_options            # dict of Engine.kde arguments
        self.kde_path = kde_path
        self.kde_format = kde_format
        self.kde_column_format = kde_column_format
        self._kde_written = False

    def run(self):
        """Advances through every interval; returns the engine's (active,
        sub-steps) after each and stops early when nothing is left to track."""
        from .engine import kde_write
        log = []
        loaded = None
        for iv in self.intervals:
            if self.flow_for_step is not None and iv.ktime != loaded:
                P, extras = self.flow_for_step(iv.ktime)
                self.E.set_flow(P)
                if extras:
                    self.E.set_cell_extras(**extras)
                # TrackCell%SteadyState follows the loaded step
                # (ParticleTrackingEngine.f90:893, 1091)
                self.E.set_steady_state(getattr(P, "steady_state", 1))
                loaded = iv.ktime
            active, steps = self.E.advance(iv.time)
            st = self.E.particles()["status"]
            pending = int((st == 0).sum())
            if iv.time_point is not None:
                if self.timeseries_path is not None:
                    self.E.write_timeseries(self.timeseries_path, iv.time_point, iv.ktime,
                                            iv.time, self.sequence, self.ids, self.group)
                for path, cells in self.observations:
                    self.E.write_observation(path, 1, cells, iv.time_point, iv.ktime, self.ids,
                                             self.group, self.group_solute, mass=self.mass)
            if iv.gpkde_point is not None and self.kde_options is not None:
                k = dict(self.kde_options)
                dens, hist = self.E.kde(mass=self.mass, **k)
                if self.kde_path is not None:
                    kde_write(self.kde_path, k["origin"], k["bin_size"], k["nbins"], dens, hist,
                              time_id=iv.gpkde_point, time=iv.time, group_id=1,
                              fmt=self.kde_format, column_format=self.kde_column_format,
                              histogram_scale=k.get("scaling_factor", 1.0),
                              append=self._kde_written)
                    self._kde_written = True
            log.append((iv, active, steps))
            if active == 0 and pending == 0:
                break
        return log
