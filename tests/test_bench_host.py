"""Host checks of bench.py's interval schedule (no GPU)."""

import itertools

import pytest

import bench


def _graph_keys(lengths, start):
    """(length, parity at its first substep) of each interval: the graph cache key's varying
    part, as Solver.prepare keys its captured graphs."""
    keys, clock = [], start
    for length in lengths:
        keys.append((length, clock % 2))
        clock += length
    return keys, clock


@pytest.mark.parametrize("steps,warmup", [(20, 5), (1, 0), (2000, 40), (41, 0), (40, 0), (80, 100), (3, 7),
                                          (39, 1), (79, 200), (20, 50)])
def test_schedule_times_exactly_the_requested_substeps(steps, warmup):
    warm, timed = bench.interval_schedule(steps, warmup)
    assert sum(timed) == steps
    assert sum(warm) >= warmup
    assert all(0 < length <= bench.INTERVAL_SUBSTEPS for length in warm + timed)
    # Every graph the timed region replays was captured during the warm-up.
    captured, clock = _graph_keys(warm, 0)
    replayed, _ = _graph_keys(timed, clock)
    assert set(replayed) <= set(captured)


def test_schedule_over_a_range_of_driver_arguments():
    for steps, warmup in itertools.product(range(1, 130), (0, 1, 5, 40, 41)):
        warm, timed = bench.interval_schedule(steps, warmup)
        captured, clock = _graph_keys(warm, 0)
        replayed, _ = _graph_keys(timed, clock)
        assert sum(timed) == steps and sum(warm) >= warmup and set(replayed) <= set(captured)


def test_schedule_rejects_empty_runs():
    with pytest.raises(ValueError):
        bench.interval_schedule(0, 5)
    with pytest.raises(ValueError):
        bench.interval_schedule(10, -1)
