"""The allocation-free maxima entry point (``mm_local_maxima_noalloc``) against the entry point with a
caller workspace (``mm_local_maxima``), through the C ABI: the same counts and raster-ordered
candidates, bit for bit, on consecutive calls (the second on a smaller volume, over the stale contents
the first left in the library's scratch array)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _both(L, img, thr, delta, cap, sign=1.0):
    dev = img.device
    D, H, W = img.shape
    st = L.stream_ptr(dev)
    ws_bytes = L.lib().mm_maxima_ws_bytes(D, H, W, 1)
    ws = torch.full((ws_bytes,), 0x5A, dtype=torch.uint8, device=dev)
    out = []
    for name in ("mm_local_maxima", "mm_local_maxima_noalloc"):
        coords = torch.full((cap, 3), -7, dtype=torch.int32, device=dev)
        count = torch.full((1,), -1, dtype=torch.int32, device=dev)
        args = (L.ptr(img), D, H, W, img.stride(0), img.stride(1), float(sign), float(thr), int(delta),
                L.ptr(coords), cap, L.ptr(count))
        if name == "mm_local_maxima":
            L.call(name, *args, L.ptr(ws), ws_bytes, st)
        else:
            L.call(name, *args, st)
        torch.cuda.synchronize(dev)
        n = int(count.item())
        out.append((n, coords[:min(n, cap)].cpu().numpy()))
    return out


@pytest.mark.parametrize("delta", [1, 3, 5])
def test_noalloc_maxima_equals_workspace_entry(dev, delta):
    from mermake_amd import _lib as L
    g = torch.Generator(device=dev)
    g.manual_seed(17 + delta)
    calls = [((40, 1800, 1900), 0.97, 1 << 17, 1.0),   # 23 MB of scratch
             ((20, 300, 257), 0.9, 1 << 16, 1.0),       # smaller, over the stale contents of the first
             ((7, 129, 65), -0.95, 1 << 12, -1.0)]      # minus sign, smaller still
    for shape, thr, cap, sign in calls:
        print("noalloc maxima", shape, "delta", delta, "sign", sign, flush=True)
        img = torch.rand(shape, generator=g, device=dev, dtype=torch.float32)
        (n1, c1), (n2, c2) = _both(L, img, thr, delta, cap, sign)
        assert n1 == n2 and n1 > 0, (shape, n1, n2)
        assert np.array_equal(c1, c2), shape


def test_noalloc_maxima_cap_and_views(dev):
    """A count above ``cap`` is reported with the first ``cap`` candidates, as the workspace entry; strided
    views (a window of a larger volume) read through their strides."""
    from mermake_amd import _lib as L
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    big = torch.rand((16, 400, 420), generator=g, device=dev, dtype=torch.float32)
    view = big[2:14, 30:330, 17:400]
    (n1, c1), (n2, c2) = _both(L, view, 0.5, 1, 1000)
    assert n1 == n2 > 1000 and np.array_equal(c1, c2) and len(c1) == 1000


def test_noalloc_maxima_rejects_oversized_volume(dev):
    from mermake_amd import _lib as L
    lib = L.lib()
    D, H, W = 48, 3000, 3000
    assert lib.mm_maxima_ws_bytes(D, H, W, 1) > 48 << 20
    one = ctypes.c_void_p(1)
    rc = lib.mm_local_maxima_noalloc(one, D, H, W, H * W, W, 1.0, 0.5, 1, one, 8, one, L.stream_ptr(dev))
    assert rc == 1001 and b"scratch" in lib.mm_last_error()
