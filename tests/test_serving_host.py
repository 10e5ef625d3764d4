"""Host-side logic of the serving loop: queue-size classes and arrivals
(no device calls)."""
import numpy as np
import pytest

from artea_amd.serving import QUEUE_CLASSES, percentile_nearest_rank, poisson_arrivals, queue_class, queue_classes


def test_queue_classes_match_scalar():
    q = np.arange(0, QUEUE_CLASSES[-1] + 1)
    assert np.array_equal(queue_classes(q), [queue_class(int(x)) for x in q])
    # every class is the smallest one that holds the queue
    for c_prev, c in zip((0,) + QUEUE_CLASSES[:-1], QUEUE_CLASSES):
        assert queue_class(c_prev + 1) == c and queue_class(c) == c
    assert queue_classes(np.array([], np.uint32)).size == 0


def test_queue_classes_reject_oversized():
    with pytest.raises(ValueError):
        queue_class(QUEUE_CLASSES[-1] + 1)
    with pytest.raises(ValueError):
        queue_classes(np.array([10, QUEUE_CLASSES[-1] + 1]))


def test_arrivals_and_percentile():
    a = poisson_arrivals(20000, 1000.0, seed=3)
    assert np.all(np.diff(a) >= 0)
    assert abs(20000 / a[-1] - 1000.0) < 50.0  # mean rate within 5%
    x = np.arange(1, 101, dtype=np.float64)
    assert percentile_nearest_rank(x, 99) == 99.0 and percentile_nearest_rank(x, 50) == 50.0
    assert percentile_nearest_rank(np.array([7.0]), 99) == 7.0
