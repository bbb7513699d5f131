"""Problem scenarios shared by the parity tests: each entry is
(problem, release cell or None for the block release, output interval,
number of output intervals)."""
import numpy as np

from rwpt import problems


def hetero(**kw):
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def closed_cell(P, k, i, j, q=0.5, inward=True):
    """Makes the face flows of cell (k, i, j) (1-based, interior) all point
    into the cell (inward: no exit face in forward tracking, a sink) or all
    out of it (no exit face in backward tracking, a source), with the
    MODFLOW-2005 signs: frf = flow from column j to j+1, fff = row i to
    i+1, flf = layer k to k+1."""
    sg = 1.0 if inward else -1.0
    c = P.cell_number(k, i, j) - 1
    P.frf[c - 1] = sg * q               # left neighbour -> cell
    P.frf[c] = -sg * q                  # right neighbour -> cell
    P.fff[c - P.ncol] = sg * q          # row above -> cell
    P.fff[c] = -sg * q                  # row below -> cell
    P.flf[c - P.ncol * P.nrow] = sg * q  # layer above -> cell
    P.flf[c] = -sg * q                  # layer below -> cell
    return c


def stopping(kind, **kw):
    """The stopping branches of TrackCell.f90:433-497 on the heterogeneous
    field: stop zones, weak sinks (forward) and weak sources (backward),
    steady or transient flow, and cells with no exit face whose sink or
    source flow is zero or not."""
    P = hetero(**kw)
    P.frf, P.fff, P.flf = P.frf.copy(), P.fff.copy(), P.flf.copy()
    g = np.random.default_rng(11)
    P.zone = np.ones(P.ncell, np.int32)
    P.sink_flow = np.zeros(P.ncell)
    P.source_flow = np.zeros(P.ncell)
    if kind == "zone":
        col = (np.arange(P.ncell) % P.ncol) + 1
        P.zone[(col >= 11) & (col <= 12)] = 3
        P.zone[g.random(P.ncell) < 0.03] = 2
        P.stop_zone = 3
        return P
    backward = bool(P.backward)
    flows = P.source_flow if backward else P.sink_flow
    weak = g.random(P.ncell) < 0.04
    flows[weak] = 0.02 if backward else -0.02
    # cells without an exit face: every other one with a weak flow
    cols = (4, 5, 6) if backward else (9, 10, 11, 12)
    m = 0
    for j in cols:
        for i in range(7, 15, 2):
            for k in (2, 4, 5):
                c = closed_cell(P, k, i, j, inward=not backward)
                flows[c] = (0.05 if backward else -0.05) if m % 2 else 0.0
                m += 1
    return P


def _scenarios():
    s = {}
    s["cfg0"] = (problems.config0(), (1, 25, 10), 0.25, 12)
    s["cfg0_reflect"] = (problems.config0(default_icbound=1), (1, 2, 97), 0.25, 12)
    s["hetero3d"] = (hetero(), None, 0.5, 8)
    P = hetero(default_icbound=1)
    P.ibound = np.ones(P.ncell, np.int32)
    P.ibound[P.cell_number(3, 10, 12) - 1] = 0
    P.ibound[P.cell_number(2, 8, 9) - 1] = 0
    P.icbound = np.zeros(P.ncell, np.int32)
    P.icbound[P.cell_number(4, 12, 6) - 1] = 1
    P.retardation = np.random.default_rng(2).uniform(1.0, 1.5, P.ncell)
    s["hetero_ibound_rebound_R"] = (P, None, 0.5, 8)
    # axisymmetric model with aLV > aLH and aTV < aTH: the weights of
    # TrackSubCell.F90:792-797 stay positive
    P = hetero(disp_model=1)
    P.aLV, P.aTV = np.array([0.8]), np.array([0.02])
    s["axisymmetric"] = (P, None, 0.5, 8)
    # the reference's cosinesq = vz^2/|v| (not vz^2/|v|^2) exceeds 1 and
    # drives alphaL negative here: sqrt of a negative gives NaN positions on
    # both sides, and the NaNs must land on the same particles
    P = hetero(disp_model=1)
    P.aLV, P.aTV = np.array([0.2]), np.array([0.01])
    s["axisymmetric_nan"] = (P, None, 0.5, 8)
    P = hetero(time_step_kind=3, ts_params=(0.3, 0.2))
    s["min_adv_disp"] = (P, None, 1.0, 6)
    P = hetero(time_step_kind=1, ts_params=(0.5, 0.0), dim_mask=(1, 0, 1))
    s["adv_xz_2d"] = (P, None, 1.0, 6)
    P = hetero(dim_mask=(0, 1, 0))
    s["1d_y"] = (P, None, 0.5, 6)
    P = hetero()
    P.aL = np.random.default_rng(4).uniform(0.1, 1.0, P.ncell)
    P.aT = P.aL * 0.1
    P.dmeff = np.full(P.ncell, 1e-4)
    s["distributed_dispersivity"] = (P, None, 0.5, 6)
    P = hetero(backward=1)
    s["backward"] = (P, None, 0.5, 6)
    # non-uniform spacings: powers of two take the exact-reciprocal path in
    # the device code, the others the IEEE division path
    P = hetero()
    g = np.random.default_rng(8)
    P.delr = g.choice([0.5, 1.0, 1.5, 2.0, 0.7], P.ncol)
    P.delc = g.choice([0.25, 1.0, 1.3], P.nrow)
    dz = g.choice([1.0, 2.0, 0.6], P.nlay)
    tops = np.concatenate([[dz.sum()], dz.sum() - np.cumsum(dz)[:-1]])
    P.top = np.repeat(tops, P.nrow * P.ncol)
    P.bot = np.repeat(tops - dz, P.nrow * P.ncol)
    s["nonuniform_spacing"] = (P, None, 0.5, 6)
    # stopping conditions of TrackCell (statuses 3, 4, 5, 6)
    s["stop_zone"] = (stopping("zone"), None, 0.5, 8)
    s["weak_sinks"] = (stopping("sink", stop_weak_sinks=1), None, 0.5, 8)
    s["weak_sources_backward"] = (stopping("source", backward=1, stop_weak_sources=1),
                                  None, 0.5, 6)
    s["transient_weak_sinks"] = (stopping("sink", stop_weak_sinks=1, steady_state=0),
                                 None, 0.5, 8)
    s["transient_no_exit"] = (stopping("sink", steady_state=0), None, 0.5, 8)
    s["transient_weak_sources_backward"] = (
        stopping("source", backward=1, stop_weak_sources=1, steady_state=0), None, 0.5, 6)
    return s


SCEN = _scenarios()

# scenarios in which some particles reach the reference's endless zero-time
# exchange between two cells (the oracle and the device stop them with
# status 9; the reference itself never returns for them)
STALLING = ("adv_xz_2d", "1d_y")


def extra_scenarios():
    """Further option combinations, checked against the compiled reference."""
    s = {}
    base = problems.config0()
    s["cfg0_1d_x"] = (problems.Problem(**{**base.__dict__, "dim_mask": (1, 0, 0)}),
                      (1, 25, 10), 0.25, 12)
    s["cfg0_1d_y"] = (problems.Problem(**{**base.__dict__, "dim_mask": (0, 1, 0)}),
                      (1, 25, 10), 0.25, 12)
    s["hetero_yz"] = (hetero(dim_mask=(0, 1, 1)), None, 0.5, 6)
    s["hetero_adv_ts"] = (hetero(time_step_kind=1, ts_params=(0.5, 0.0)), None, 1.0, 4)
    s["hetero_disp_ts"] = (hetero(time_step_kind=2, ts_params=(0.5, 0.0)), None, 1.0, 4)
    return s


def release(P, where, n, seed=5):
    if where is None:
        return problems.release_in_block(P, n, (3, 8), (6, 15), (1, P.nlay), seed=seed)
    return problems.release_in_cell(P, n, *where, seed=seed, lz=0.5)
