"""bench.py's command line: the invocations the benchmark runs under."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_round_end_command_is_accepted():
    a = bench.parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    assert (a.gpus, a.steps, a.warmup) == (1, 20, 5)
    assert a.particles == 10_000_000 and a.precision == 1


@pytest.mark.parametrize("steps", [1, 7, 20, 21, 45, 200])
def test_any_step_count_is_accepted(steps):
    assert bench.parse_args(["--steps", str(steps)]).steps == steps


def test_defaults_are_the_full_run():
    a = bench.parse_args([])
    assert a.gpus == 1 and a.steps == bench.RUN_INTERVALS
    assert bench.RUN_INTERVALS * bench.INTERVAL_STEPS == 1000


@pytest.mark.parametrize("argv", [["--gpus", "2"], ["--gpus", "8"], ["--steps", "0"],
                                  ["--adjoint-seeds", "1"]])
def test_unmeetable_requests_exit_2(argv):
    with pytest.raises(SystemExit) as ex:
        bench.parse_args(argv)
    assert ex.value.code == 2


def test_quick_drops_the_side_legs():
    a = bench.parse_args(["--quick"])
    assert a.ensemble_intervals == 0 and a.no_adjoint and a.kde_loops == 0
