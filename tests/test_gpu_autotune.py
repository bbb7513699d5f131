"""The launch-shape autotuner (rwpt_autotune): it leaves the particle state
and engine time as they were, picks a shape from rwpt_set_tuning's space,
and the walk that follows is bit-identical to one never tuned."""
import numpy as np
import pytest

from rwpt import Engine, problems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_state_and_results(precision):
    P = problems.heterogeneous_flow(24, 20, 6, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    rel = problems.release_in_block(P, 5000, (2, 6), (6, 15), (1, 6), seed=3)
    print(f"autotune precision {precision}", flush=True)
    A, B = Engine(P, precision=precision), Engine(P, precision=precision)
    try:
        for E in (A, B):
            E.load_particles(*rel)
            E.advance(2.0)
        before = A.particles()
        best = A.autotune(6.0)
        after = A.particles()
        for k in before:
            np.testing.assert_array_equal(before[k], after[k])
        assert best["block"] in (64, 128, 256) and best["min_waves"] in (1, 3, 4)
        assert best["budget"] in (16, 32, 64, 128, 256) and best["sort"] in (0, 1)
        assert best["ms"] > 0
        for E in (A, B):
            E.advance(6.0)
            E.advance(10.0)
        a, b = A.particles(), B.particles()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
    finally:
        A.close()
        B.close()
