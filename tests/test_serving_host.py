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


def test_oversized_request_is_refused():
    """A request larger than a slot (N > n_max) would overwrite the next slot's state; it
    is refused before anything is admitted."""
    import pytest
    from packing_amd.serving import Request, check_request
    ok = Request(0, np.zeros((40, 3), np.float32), np.ones(3, np.float32))
    check_request(ok, 40)
    with pytest.raises(ValueError):
        check_request(Request(1, np.zeros((41, 3), np.float32), np.ones(3, np.float32)), 40)
    with pytest.raises(ValueError):
        check_request(Request(2, np.zeros((0, 3), np.float32), np.ones(3, np.float32)), 40)
    with pytest.raises(ValueError):
        check_request(Request(3, np.zeros((10, 2), np.float32), np.ones(3, np.float32)), 40)


def test_size_classes():
    """Size classes follow the one-wave kernel's register rows R = ceil((3N + 3) / 64):
    R <= 3, 4-5, 6-8, 9-10, then the workgroup kernel."""
    from packing_amd.serving import size_class
    assert [size_class(n) for n in (1, 63, 64, 105, 106, 169, 170, 212, 213, 400)] == \
        [0, 0, 1, 1, 2, 2, 3, 3, 4, 4]
