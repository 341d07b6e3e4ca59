"""Host-side pieces of the request loop (no GPU)."""
import numpy as np


def test_poisson_arrivals():
    from mermake_amd.serve import poisson_arrivals
    a = poisson_arrivals(2000, 4.0, seed=3)
    assert a.shape == (2000,) and np.all(np.diff(a) >= 0) and a[0] >= 0
    assert abs(len(a) / a[-1] - 4.0) < 0.4          # mean rate
    assert np.array_equal(poisson_arrivals(5, 0), np.zeros(5))
    assert np.array_equal(poisson_arrivals(10, 2.0, seed=1), poisson_arrivals(10, 2.0, seed=1))


def test_request_latency():
    from mermake_amd.serve import FovRequest
    r = FovRequest(7, [np.zeros((2, 3, 3), np.uint16)], arrival=1.5)
    assert r.latency is None and r.rid == 7 and len(r.channels) == 1
    r.t_done = 2.0
    assert r.latency == 0.5


def test_request_channel_flats():
    import pytest
    from mermake_amd.serve import FovRequest
    a, b = np.zeros((3, 3), np.float32), np.ones((3, 3), np.float32)
    chans = [np.zeros((2, 3, 3), np.uint16)] * 2
    assert FovRequest(0, chans, a).channel_flats() == [a, a]
    assert FovRequest(0, chans, [a, b]).channel_flats() == [a, b]
    assert FovRequest(0, chans, None).channel_flats() == [None, None]
    with pytest.raises(ValueError):
        FovRequest(0, chans, [a]).channel_flats()
