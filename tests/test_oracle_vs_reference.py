"""The oracle against the reference itself.

oracle/ref/Makefile compiles the reference's Fortran tracking engine
(ParticleTrackingEngine.f90 with TrackCell, TrackSubCell, ModpathCellData,
rng_par_zig and the grid classes) with amdflang, and oracle/ref/rwpt_ref.f90
drives it with the particle loop of MPathRW.F90:1239-1403.  The oracle must
reproduce it bit for bit: cell, status, local coordinates and time after
every output interval, NaNs included.

In the two scenarios where some particles fall into the reference's endless
zero-time exchange between two cells, the oracle stops those particles with
status 9 and the reference never returns; they are marked as already
terminated (status 2) on both sides and every other particle is compared.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import refrun
from rwpt import problems
from scenarios import SCEN, STALLING, extra_scenarios, release

pytestmark = pytest.mark.skipif(not refrun.available(),
                                reason="reference driver not built (oracle/ref)")

ALL = {**SCEN, **extra_scenarios()}
N = 200


def _oracle_run(P, cell, lx, ly, lz, dt_out, nout, t=None, status=None):
    st = O.new_state(cell, lx, ly, lz)
    if t is not None:
        st["t"] = np.array(t, dtype=np.float64)
    if status is not None:
        st["status"] = np.array(status, dtype=np.int32)
    OP = O.OracleProblem(P)
    out, t0 = [], 0.0
    for k in range(1, nout + 1):
        O.advance(OP, st, t0, k * dt_out)
        t0 = k * dt_out
        out.append({f: st[f].copy() for f in ("cell", "status", "lx", "ly", "lz", "t")})
    return out


def _same(a, b):
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


@pytest.mark.parametrize("name", sorted(ALL))
@pytest.mark.parametrize("advection", [2, 1])
def test_oracle_matches_reference(name, advection):
    P, where, dt_out, nout = ALL[name]
    if advection != P.advection_kind:
        if name not in ("cfg0", "hetero3d", "hetero_ibound_rebound_R", "backward"):
            pytest.skip("exponential advection is checked on four scenarios")
        P = problems.Problem(**{**P.__dict__, "advection_kind": advection})
    cell, lx, ly, lz = release(P, where, N)
    status = np.ones(N, np.int32)
    if name in STALLING:
        first = _oracle_run(P, cell, lx, ly, lz, dt_out, nout)
        stalled = first[-1]["status"] == 9
        assert 0 < stalled.sum() <= N // 20
        status[stalled] = 2
    want = _oracle_run(P, cell, lx, ly, lz, dt_out, nout, status=status)
    got = refrun.run(P, cell, lx, ly, lz, [k * dt_out for k in range(1, nout + 1)],
                     status=status, timeout=300)
    for k in range(nout):
        for f in ("cell", "status", "lx", "ly", "lz", "t"):
            ok = _same(got[k][f], want[k][f])
            assert ok.all(), (k, f, int((~ok).sum()))
    assert (want[-1]["status"] != 2).sum() > 0


def test_pending_release_matches_reference():
    """Release times before, inside and after the intervals: statuses 0, 1
    and 8 as MPathRW.F90:1259-1268 assigns them."""
    P, where, _, _ = SCEN["hetero3d"]
    cell, lx, ly, lz = release(P, where, N)
    t0 = np.random.default_rng(1).uniform(-1.0, 6.0, N)
    status = np.zeros(N, np.int32)
    want = _oracle_run(P, cell, lx, ly, lz, 2.0, 4, t=t0, status=status)
    got = refrun.run(P, cell, lx, ly, lz, [2.0, 4.0, 6.0, 8.0], t=t0, status=status)
    for k in range(4):
        for f in ("cell", "status", "lx", "ly", "lz", "t"):
            assert np.array_equal(got[k][f], want[k][f]), (k, f)
    assert set(np.unique(want[0]["status"])) >= {0, 1, 8}
