"""Graph mode of the advance (rwpt_set_graph): each interval's walk runs as
replays of a captured HIP graph of K walk launches.  The results must be
bit-identical to the eager walk, which the oracle pins: every interval's
state after the replays (cell, status, positions, time, sub-steps, rng),
with eager work between the replays (prepare launches, KDE reconstructions,
host reads) and the graph re-captured and updated for every advance.
Structured FP64 and FP32-local engines, an ensemble and an unstructured
grid."""
import numpy as np
import pytest

from oracle import oracle as O
from rwpt import Engine, problems
from scenarios import SCEN, release as _release

pytestmark = pytest.mark.gpu

FIELDS = ("cell", "status", "lx", "ly", "lz", "t", "substeps", "rng")


def _advance_all(P, rel, times, graph, precision=0, kde=None, budget=16, max_list=0):
    E = Engine(P, precision=precision)
    out, launches = [], []
    try:
        E.set_tuning(budget=budget, block=256, sort=True)
        E.set_graph(graph, max_list)
        E.load_particles(*rel)
        for t in times:
            E.advance(t)
            launches.append(E.last_launches())
            out.append(E.particles())
            if kde is not None:
                E.kde(**kde)
        ms, nl = E.kernel_time(0)
    finally:
        E.close()
    return out, launches, nl


@pytest.mark.parametrize("graph,max_list", [(2, 0), (8, 0), (4, 300)])
def test_graph_equals_eager_and_oracle(graph, max_list):
    P, where, dt_out, nout = SCEN["hetero_ibound_rebound_R"]
    rel = _release(P, where, 1500)
    times = [k * dt_out for k in range(1, 4)]
    kde = dict(origin=(0.0, 0.0, float(P.bot.min())), bin_size=(1.0, 1.0, 1.0),
               nbins=(P.ncol, P.nrow, P.nlay), mode=1)
    print("graph: eager", flush=True)
    ref, l0, _ = _advance_all(P, rel, times, 0, kde=kde)
    print("graph: replays", graph, flush=True)
    got, l1, nl = _advance_all(P, rel, times, graph, kde=kde, max_list=max_list)
    st = O.new_state(*rel)
    OP = O.OracleProblem(P)
    t0 = 0.0
    for k, (a, b) in enumerate(zip(ref, got)):
        O.advance(OP, st, t0, times[k])
        t0 = times[k]
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), (k, f)
            assert np.array_equal(b[f], st[f]), (k, f)
    assert nl == sum(l1)
    if max_list:
        # eager cell-sorted launches first, then whole graphs
        assert all(n > graph for n in l1)
        return
    # whole graphs of K launches (with K = 2 several replays an interval)
    assert all(n % graph == 0 and n >= graph for n in l1)
    assert graph > 2 or max(l1) > graph


def test_graph_fp32_and_ensemble():
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    rel = problems.release_in_block(P, 3000, (2, 8), (4, 16), (1, 6), seed=3)
    times = [1.0, 2.5, 4.0]
    for prec in (0, 1):
        print("graph: precision", prec, flush=True)
        ref, _, _ = _advance_all(P, rel, times, 0, precision=prec)
        got, _, _ = _advance_all(P, rel, times, 4, precision=prec)
        for a, b in zip(ref, got):
            for f in FIELDS:
                assert np.array_equal(a[f], b[f], equal_nan=f in ("lx", "ly", "lz", "t")), f
    members = []
    for k in range(3):
        Q = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
        Q.aL, Q.aT = np.array([0.2 + 0.3 * k]), np.array([0.02 + 0.01 * k])
        Q.dmeff, Q.ts_params = np.array([1e-4]), (0.5, 0.0)
        members.append(Q)
    S = problems.stack(members)
    cells = np.concatenate([rel[0] + k * P.ncell for k in range(3)]).astype(np.int32)
    rel3 = (cells,) + tuple(np.concatenate([r] * 3) for r in rel[1:])
    print("graph: ensemble", flush=True)
    ref, _, _ = _advance_all(S, rel3, times, 0)
    got, _, _ = _advance_all(S, rel3, times, 6)
    for a, b in zip(ref, got):
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), f


def test_graph_unstructured_grid():
    from usg_scenarios import build, release
    Q, U = build("checker")
    rel = release(U, 800, seed=3)
    times = [0.3, 0.7, 1.5]
    print("graph: unstructured", flush=True)
    ref, _, _ = _advance_all(U, rel, times, 0, budget=4)
    got, _, _ = _advance_all(U, rel, times, 4, budget=4)
    for a, b in zip(ref, got):
        for f in FIELDS:
            assert np.array_equal(a[f], b[f], equal_nan=f in ("lx", "ly", "lz", "t")), f


def test_set_graph_rejects_odd_counts():
    P, where, dt_out, nout = SCEN["hetero_ibound_rebound_R"]
    E = Engine(P)
    try:
        with pytest.raises(Exception):
            E.set_graph(3)
        E.set_graph(0)
    finally:
        E.close()
