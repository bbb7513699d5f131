"""The tracking-interval schedule of MPathRW (rwpt.driver.schedule), with
known answers worked out by hand from TIME_STEP_LOOP and
TRACKING_INTERVAL_LOOP (MPathRW.F90:850-1733, :1119-1186)."""
from rwpt.driver import schedule


def times(iv):
    return [(i.ktime, i.time, i.time_point, i.gpkde_point, i.step_end) for i in iv]


def test_timeseries_points_in_one_extended_step():
    # steady state: one step extended to the stop time (StoppingTimeOption 2)
    iv = schedule([1e30], 10.0, time_points=[2.0, 4.0, 6.0])
    assert times(iv) == [(1, 2.0, 1, None, False), (1, 4.0, 2, None, False),
                         (1, 6.0, 3, None, False), (1, 10.0, None, None, True)]


def test_gpkde_following_the_timeseries():
    iv = schedule([1e30], 10.0, time_points=[2.0, 4.0, 6.0], reconstruct=True)
    assert [(i.time, i.time_point, i.gpkde_point) for i in iv] == \
        [(2.0, 1, 1), (4.0, 2, 2), (6.0, 3, 3), (10.0, None, None)]


def test_gpkde_times_of_their_own():
    # a reconstruction time before the pending timeseries point ends the
    # interval there, and that timeseries point is not written (its index
    # was consumed, :1170-1176); equal times share the interval
    iv = schedule([1e30], 10.0, time_points=[2.0, 4.0, 6.0], gpkde_points=[3.0, 4.0, 8.0],
                  gpkde_follows_timeseries=False, reconstruct=True)
    assert times(iv) == [(1, 2.0, 1, None, False), (1, 3.0, None, 1, False),
                         (1, 4.0, None, 2, False), (1, 8.0, None, 3, False),
                         (1, 10.0, None, None, True)]
    iv = schedule([1e30], 10.0, time_points=[2.0, 4.0], gpkde_points=[4.0],
                  gpkde_follows_timeseries=False, reconstruct=True)
    assert [(i.time, i.time_point, i.gpkde_point) for i in iv] == \
        [(2.0, 1, None), (4.0, 2, 1), (10.0, None, None)]


def test_several_time_steps():
    # StoppingTimeOption 1: every step ends at its own end time; a point
    # beyond the step waits for the next one
    iv = schedule([5.0, 10.0], 10.0, time_points=[3.0, 7.0], stopping_option=1)
    assert times(iv) == [(1, 3.0, 1, None, False), (1, 5.0, None, None, True),
                         (2, 7.0, 2, None, False), (2, 10.0, None, None, True)]
    # StoppingTimeOption 3: a step beyond the stop time is cut there and
    # tracking ends
    iv = schedule([5.0, 10.0, 15.0], 8.0, time_points=[3.0], stopping_option=3)
    assert times(iv) == [(1, 3.0, 1, None, False), (1, 5.0, None, None, True),
                         (2, 8.0, None, None, True)]


def test_point_at_a_step_end_closes_the_step():
    iv = schedule([5.0, 10.0], 10.0, time_points=[5.0, 10.0], stopping_option=1)
    assert times(iv) == [(1, 5.0, 1, None, True), (2, 10.0, 2, None, True)]
