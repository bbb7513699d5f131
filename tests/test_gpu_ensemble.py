"""Ensemble engines (rwpt_create_ensemble): members stacked in one engine
walk in the same launches, and every member's particles end exactly where
the same member run alone puts them (bit-exact, both precisions), with the
member's rng streams taken at its offset.  Members differ in porosity
(uniform and non-uniform), dispersivities, retardation and particle count."""
import numpy as np
import pytest

from rwpt import Engine, problems

pytestmark = pytest.mark.gpu


def members():
    base = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    base.ts_params = (0.5, 0.0)
    out = []
    rng = np.random.default_rng(3)
    for m in range(4):
        P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
        P.ts_params = (0.5, 0.0)
        if m == 1:
            P.porosity = rng.uniform(0.2, 0.35, P.ncell)      # non-uniform
        else:
            P.porosity = np.full(P.ncell, rng.uniform(0.2, 0.35))
        if m == 2:
            P.retardation = np.full(P.ncell, 2.0)
        aL = rng.uniform(0.05, 1.0)
        P.aL = np.array([aL])
        P.aT = np.array([aL * rng.uniform(0.05, 0.2)])
        P.dmeff = np.array([1e-4])
        out.append(P)
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_equals_members_alone(precision):
    M = members()
    counts = [700, 2500, 150, 1300]
    rel = [problems.release_in_block(P, n, (2, 6), (6, 15), (1, 6), seed=10 + m)
           for m, (P, n) in enumerate(zip(M, counts))]
    E = problems.stack(M)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    cell = np.concatenate([E.global_cell(m, r[0]) for m, r in enumerate(rel)]).astype(np.int32)
    lx, ly, lz = (np.concatenate([r[k] for r in rel]) for k in (1, 2, 3))
    print(f"ensemble precision {precision}: members {len(M)} particles {sum(counts)}",
          flush=True)
    eng = Engine(E, precision=precision)
    try:
        eng.load_particles(cell, lx, ly, lz)
        for k in (1, 2):
            eng.advance(8.0 * k)
        got = eng.particles()
        sub = got["substeps"]
        kde = [eng.kde((0.0, 0.0, float(M[0].bot.min())), (1.0, 1.0, 1.0), (24, 20, 6),
                       mode=1, member=m)[0] for m in range(len(M))]
    finally:
        eng.close()
    for m, P in enumerate(M):
        one = Engine(P, precision=precision)
        try:
            one.load_particles(*rel[m], stream_offset=int(offs[m]))
            for k in (1, 2):
                one.advance(8.0 * k)
            want = one.particles()
            wsub = want["substeps"]
            wkde = one.kde((0.0, 0.0, float(P.bot.min())), (1.0, 1.0, 1.0), (24, 20, 6),
                           mode=1)[0]
        finally:
            one.close()
        s = slice(offs[m], offs[m] + counts[m])
        np.testing.assert_array_equal(got["cell"][s] - m * E.ncell_member, want["cell"])
        np.testing.assert_array_equal(got["status"][s], want["status"])
        for f in ("lx", "ly", "lz", "t"):
            np.testing.assert_array_equal(got[f][s], want[f])
        np.testing.assert_array_equal(got["rng"][2 * offs[m]:2 * (offs[m] + counts[m])],
                                      want["rng"])
        np.testing.assert_array_equal(sub[s], wsub)
        # the member's reconstruction: same histogram, bandwidth statistics
        # summed in another order
        np.testing.assert_allclose(kde[m], wkde, rtol=0, atol=1e-12 * wkde.max())
