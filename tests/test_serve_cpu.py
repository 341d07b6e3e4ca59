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


def test_arena_ring_blocks_never_overlap():
    """The input arena's ring: random admissions and out-of-order completions never hand out
    overlapping or out-of-range blocks, a block is refused only when it does not fit the free room, and
    an emptied ring starts again at 0."""
    import random
    from mermake_amd.serve import _Arena
    A = _Arena(64 * 256, "cpu")
    assert A.size == 64 * 256 and A.alloc(A.size + 256) is None
    rng = random.Random(5)
    live, refused = [], 0
    for _ in range(20000):
        if live and rng.random() < 0.45:
            A.free(live.pop(rng.randrange(len(live))))
        else:
            n = 256 * rng.choice([1, 2, 4, 6, 8, 13])
            b = A.alloc(n)
            if b is None:
                refused += 1
                assert live        # an empty ring takes any block up to its size
                continue
            assert 0 <= b[0] < b[1] <= A.size and b[1] - b[0] == n and b[0] % 256 == 0
            assert all(b[1] <= o[0] or o[1] <= b[0] for o in live)
            live.append(b)
    assert refused > 0
    for b in live:
        A.free(b)
    assert not A.blocks and A.alloc(A.size) == [0, A.size, False]


def test_arena_views_keep_dtype_shape_and_tensor_type():
    import torch
    from mermake_amd.deconvolver import SquaredU8
    from mermake_amd.serve import _Arena
    A = _Arena(1 << 16, "cpu")
    blk = A.alloc(_Arena.rounded(2 * 3 * 4 * 2) + _Arena.rounded(5 * 7))
    h = torch.arange(24, dtype=torch.int32).to(torch.uint16).view(2, 3, 4)
    u = SquaredU8.of(np.arange(35, dtype=np.uint8).reshape(1, 5, 7))
    v = A.view(blk, 0, h)
    v.copy_(h)
    w = A.view(blk, _Arena.rounded(h.numel() * 2), u)
    w.copy_(u)
    assert v.dtype == torch.uint16 and tuple(v.shape) == (2, 3, 4) and type(v) is torch.Tensor
    assert type(w) is SquaredU8 and tuple(w.shape) == (1, 5, 7)
    assert torch.equal(v.to(torch.int32), h.to(torch.int32)) and torch.equal(w.as_subclass(torch.Tensor), u)
