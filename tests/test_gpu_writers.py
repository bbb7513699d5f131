"""Timeseries and endpoint files written from the engine (rwpt_write_*)
against the reference's own writers (WriteTimeseriesHeader/Record,
WriteEndpoints, WriteEndpointsMassParticles of ParticleManager.F90, compiled
here into oracle/_ref/writers_ref) on the same particles: byte-identical."""
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
    print("writers: particles", n, flush=True)
    E = Engine(P)
    try:
        E.set_cell_extras(zone=zones)
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
    finally:
        E.close()
    assert {0, 1, 2, 8} <= set(np.unique(st["status"]).tolist())
    meta = {"group": group, "sequence": seq, "id": ids, "mass": mass, "solute": solute,
            "names": names, "tracking_direction": 1, "time_point": 3, "time_step": 1,
            "reference_time": 0.5, "origin": (100.25, -3.5), "rotation": 12.0,
            "time": 8.0, "zones": zones}
    init = {"cell": cell, "lx": lx, "ly": ly, "lz": lz, "t": t0, "face": face}
    ref_ts, ref_ep, ref_epm = refrun.writers(P, st, init, meta)
    assert ts.read_bytes() == ref_ts
    assert ep.read_bytes() == ref_ep
    assert epm.read_bytes() == ref_epm
