"""Timeseries, endpoint and observation files written from the engine
(rwpt_write_*) against the reference's own writers (WriteTimeseriesHeader/
Record, WriteEndpoints, WriteEndpointsMassParticles of ParticleManager.F90,
WriteResidentObsRecord[Binary] and WriteSinkObsRecord[Binary] of
Observation.f90, compiled here into oracle/_ref/writers_ref) on the same
particles: byte-identical.  The water volume of a resident record is the
porosity times DX DY and the thickness of the (confined, wet) cell, given to
the reference driver from the problem arrays.  The top layer's saturated top
(Grid%SaturatedTop) lies below its cell top, so the global z of its
particles follows ConvertToModelXYZ with the saturated top."""
import numpy as np
import pytest

from oracle import refrun
from rwpt import Engine, engine, problems

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refrun.writers_available(),
                                 reason="reference writers not built")]


def test_files_match_reference(tmp_path):
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    g = np.random.default_rng(9)
    zones = g.integers(1, 6, P.ncell).astype(np.int32)
    counts = [400, 300, 100]
    rel = [problems.release_in_block(P, n, (2, 8), (4, 16), (1, 6), seed=20 + k)
           for k, n in enumerate(counts)]
    cell, lx, ly, lz = (np.concatenate([r[k] for r in rel]) for k in range(4))
    n = cell.size
    group = np.concatenate([np.full(c, k + 1, np.int32) for k, c in enumerate(counts)])
    ids = np.concatenate([np.arange(1, c + 1, dtype=np.int32) for c in counts])
    seq = np.arange(1, n + 1, dtype=np.int32)
    t0 = np.zeros(n)
    status = np.ones(n, np.int32)
    last = slice(n - counts[2], n)
    t0[last] = 2.0                               # released during the run
    status[n - 40:n - 20] = 0                    # pending beyond the run
    t0[n - 40:n - 20] = 1e6
    status[n - 20:] = 8                          # never released
    face = g.integers(0, 7, n).astype(np.int32)
    mass = g.uniform(0.5, 2.0, n)
    # weak sinks in a band of cells across the plume's path, observation
    # cells over part of the band and of the plume
    sink = np.zeros(P.ncell)
    k, i, j = np.meshgrid(np.arange(1, 7), np.arange(6, 15), np.arange(11, 12), indexing="ij")
    sink[P.cell_number(k.ravel(), i.ravel(), j.ravel()) - 1] = -0.3
    k, i, j = np.meshgrid(np.arange(2, 6), np.arange(5, 13), np.arange(4, 12), indexing="ij")
    obs_cells = np.sort(P.cell_number(k.ravel(), i.ravel(), j.ravel())).astype(np.int32)
    flow_rate = float(sink[obs_cells - 1].sum())
    # the saturated top of a convertible top layer (the top of the global z)
    sat = P.top.copy()
    sat[:P.ncol * P.nrow] -= 0.4 * (P.top - P.bot)[:P.ncol * P.nrow]
    print("writers: particles", n, flush=True)
    E = Engine(P, stop_at_weak_sinks=1)
    try:
        E.set_cell_extras(zone=zones, sink_flow=sink)
        E.set_saturated_top(sat)
        E.load_particles(cell, lx, ly, lz, t=t0, status=status)
        E.advance(8.0)
        st = E.particles()
        ts = tmp_path / "ts.out"
        engine.write_timeseries_header(str(ts), 1, 0.5, (100.25, -3.5), 12.0)
        E.write_timeseries(str(ts), 3, 1, 8.0, seq, ids, group)
        ep, epm = tmp_path / "ep.out", tmp_path / "epm.out"
        names = ["inlet", "second group", "late"]
        solute = np.array([1, 2, 1], np.int32)
        args = (1, 0.5, (100.25, -3.5), 12.0, names, solute, group, seq, ids, cell, lx, ly,
                lz, t0)
        E.write_endpoints(str(ep), 0, *args, initial_face=face)
        E.write_endpoints(str(epm), 1, *args, mass=mass)
        obs_files = {}
        for style in (1, 2):
            for fmt in (0, 1):
                f = tmp_path / f"obs{style}{fmt}"
                E.write_observation(str(f), style, obs_cells, 3, 1, ids, group, solute,
                                    mass=mass, flow_rate=flow_rate, fmt=fmt, append=False)
                obs_files[style, fmt] = f.read_bytes()
    finally:
        E.close()
    assert {0, 1, 2, 3, 8} <= set(np.unique(st["status"]).tolist())
    in_obs = np.isin(st["cell"], obs_cells)
    assert (in_obs & (st["status"] == 1)).any() and (in_obs & (st["status"] == 3)).any()
    meta = {"group": group, "sequence": seq, "id": ids, "mass": mass, "solute": solute,
            "names": names, "tracking_direction": 1, "time_point": 3, "time_step": 1,
            "reference_time": 0.5, "origin": (100.25, -3.5), "rotation": 12.0,
            "time": 8.0, "zones": zones, "obs_cells": obs_cells,
            "retardation": np.ones(P.ncell), "flow_rate": flow_rate, "saturated_top": sat,
            "water_volume": (P.porosity * P.delr[np.arange(P.ncell) % P.ncol] *
                             P.delc[(np.arange(P.ncell) // P.ncol) % P.nrow] *
                             (P.top - P.bot))}
    init = {"cell": cell, "lx": lx, "ly": ly, "lz": lz, "t": t0, "face": face}
    ref_ts, ref_ep, ref_epm, res_t, res_b, snk_t, snk_b = refrun.writers(P, st, init, meta)
    assert ts.read_bytes() == ref_ts
    assert ep.read_bytes() == ref_ep
    assert epm.read_bytes() == ref_epm
    assert len(res_t) > 0 and len(snk_t) > 0
    assert obs_files[1, 0] == res_t
    assert obs_files[1, 1] == res_b
    assert obs_files[2, 0] == snk_t
    assert obs_files[2, 1] == snk_b
