"""Host-side serving arithmetic (no GPU): slot memory accounting and request generation."""
import numpy as np

from packing_amd.serving import poisson_requests, slot_bytes


def test_slot_bytes():
    # x, g, and 2 * history vectors of 3N + 3 floats, plus the optimizer record
    assert slot_bytes(200, 8) == 4 * 603 * 18 + 184 + 16


def test_poisson_requests():
    r = poisson_requests(2000, rate=2000.0)
    t = np.array([q.arrival for q in r])
    n = np.array([q.xyt.shape[0] for q in r])
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / len(r) - 1 / 2000.0) < 0.1 / 2000.0
    assert n.min() >= 10 and n.max() <= 200
    # log-uniform: the median is near sqrt(10 * 200)
    assert 35 < np.median(n) < 55
