"""3D local maxima + delta_fit refinement (reference ``mermake/maxima.py:17-98``),
running ``mm_local_maxima`` and ``mm_delta_fit``.

Differences from the reference that callers can observe:
* rows come back in raster order of the maxima (the reference's order is that
  of its atomic counter, i.e. arbitrary); compare row sets, not row order;
* neighbours that a single reflection cannot bring inside a tiny axis
  (axis length <= delta) are skipped instead of read out of bounds.
"""
import torch

from . import _lib as L
from ._dev import as_device, unit_last_stride

_ws_cache = {}


def _workspace(device, nbytes):
    key = (device, "maxima")
    ws = _ws_cache.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=device)
        _ws_cache[key] = ws
    return ws


def raw_type(raw):
    """``raw_type`` of ``mm_delta_fit`` for a prepared raw volume: 0 uint16, 1 float32, 2 a
    ``SquaredU8`` stack (uint8 values whose squares are the pixels)."""
    from .deconvolver import SquaredU8
    if isinstance(raw, SquaredU8):
        if raw.dtype != torch.uint8:
            raise TypeError("a SquaredU8 volume must hold uint8 values")
        return 2
    if raw.dtype == torch.float32:
        return 1
    if raw.dtype == torch.uint16:
        return 0
    raise TypeError(f"raw volumes are uint16, float32 or SquaredU8, not {raw.dtype}")


def _raw_view(raw, shape):
    from .deconvolver import SquaredU8
    if not isinstance(raw, SquaredU8) and raw.dtype not in (torch.float32, torch.uint16):
        raw = raw.to(torch.float32)   # maxima.py:83-84: raw is cast to float32
    if tuple(raw.shape) != tuple(shape):
        raise ValueError(f"raw shape {tuple(raw.shape)} does not match image shape {tuple(shape)}")
    return unit_last_stride(raw)


def maxima_coords(image, threshold, delta, sign=1.0, cap=None):
    """Raster-ordered int32 (N, 3) maxima of ``sign*image`` (``maxima.cu:1-145``)."""
    D, H, W = image.shape
    nvox = D * H * W
    dev = image.device
    st = L.stream_ptr(dev)
    ws_bytes = L.lib().mm_maxima_ws_bytes(D, H, W, 1)
    ws = _workspace(dev, ws_bytes)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    cap = int(min(nvox, 1 << 16) if cap is None else cap)
    cap = max(cap, 1)
    while True:
        coords = torch.empty((cap, 3), dtype=torch.int32, device=dev)
        L.call("mm_local_maxima", L.ptr(image), D, H, W, image.stride(0), image.stride(1), float(sign),
               float(threshold), int(delta), L.ptr(coords), cap, L.ptr(count), L.ptr(ws), ws.numel(), st)
        n = int(count.item())          # host sync, as maxima.py:46-50 / 66-69
        if n <= cap:
            return coords[:n], count
        cap = n


def find_local_maxima(image, threshold=None, delta=1, delta_fit=0, raw=None, sigmaZ=1, sigmaXY=1.5, **kwargs):
    """``find_local_maxima``: ``(N, 8)`` float32 rows ``[zc, xc, yc, background, a, habs, hn, h]``."""
    if delta > 5:
        raise TypeError("Delta must be an less than or equal to 5 due to MAX_KERNEL_POINTS")
    image = as_device(image)
    if image.dtype != torch.float32:
        image = image.to(torch.float32)
    image = unit_last_stride(image)
    if image.ndim != 3:
        raise ValueError("find_local_maxima expects a 3D (z, x, y) volume")
    thr = float("nan") if threshold is None else float(torch.tensor(float(threshold), dtype=torch.float32))
    coords, count = maxima_coords(image, thr, delta)
    n = coords.shape[0]
    if n == 0:
        return torch.zeros((0, 8), dtype=torch.float32, device=image.device)
    if raw is None:
        raise AttributeError("'NoneType' object has no attribute 'dtype'")   # maxima.py:83
    raw = _raw_view(as_device(raw, image.device), image.shape)
    out = torch.zeros((n, 8), dtype=torch.float32, device=image.device)
    D, H, W = image.shape
    L.call("mm_delta_fit", L.ptr(image), D, H, W, image.stride(0), image.stride(1), 1.0, L.ptr(raw),
           raw_type(raw), raw.stride(0), raw.stride(1), L.ptr(coords), L.ptr(count), n,
           int(delta_fit), float(sigmaZ), float(sigmaXY), L.ptr(out), L.stream_ptr(image.device))
    return out


def find_local_maxima_signed(image, raw, threshold, delta, delta_fit, sigmaZ, sigmaXY, sign):
    """``find_local_maxima`` of ``sign*image`` without materialising the negated volume
    (the DAPI minus pass of ``__main__.py:224-225``).  ``raw``: a prepared stack (uint16, float32 or
    ``SquaredU8``) of the image's shape."""
    raw = _raw_view(raw, image.shape)
    coords, count = maxima_coords(image, threshold, delta, sign=sign)
    n = coords.shape[0]
    if n == 0:
        return torch.zeros((0, 8), dtype=torch.float32, device=image.device)
    out = torch.zeros((n, 8), dtype=torch.float32, device=image.device)
    D, H, W = image.shape
    L.call("mm_delta_fit", L.ptr(image), D, H, W, image.stride(0), image.stride(1), float(sign), L.ptr(raw),
           raw_type(raw), raw.stride(0), raw.stride(1), L.ptr(coords), L.ptr(count), n,
           int(delta_fit), float(sigmaZ), float(sigmaXY), L.ptr(out), L.stream_ptr(image.device))
    return out


def find_local_maxima_pm(image, raw, threshold, delta, delta_fit, sigmaZ, sigmaXY):
    """``(find_local_maxima_signed(..., sign=+1), find_local_maxima_signed(..., sign=-1))`` with the
    maxima of both signs from one pass over the volume (``mm_local_maxima_pm``) and one host
    synchronisation for both counts."""
    raw = _raw_view(raw, image.shape)
    D, H, W = image.shape
    dev = image.device
    st = L.stream_ptr(dev)
    ws_bytes = L.lib().mm_maxima_ws_bytes(D, H, W, 2)
    ws = _workspace(dev, ws_bytes)
    counts = torch.zeros(2, dtype=torch.int32, device=dev)
    cap = max(1, int(min(D * H * W, 1 << 16)))
    while True:
        coords = torch.empty((2, cap, 3), dtype=torch.int32, device=dev)
        L.call("mm_local_maxima_pm", L.ptr(image), D, H, W, image.stride(0), image.stride(1), float(threshold),
               int(delta), L.ptr(coords), cap, L.ptr(counts), L.ptr(ws), ws.numel(), st)
        n = counts.tolist()
        if max(n) <= cap:
            break
        cap = max(n)
    out = []
    for v, sign in ((0, 1.0), (1, -1.0)):
        if n[v] == 0:
            out.append(torch.zeros((0, 8), dtype=torch.float32, device=dev))
            continue
        rows = torch.zeros((n[v], 8), dtype=torch.float32, device=dev)
        L.call("mm_delta_fit", L.ptr(image), D, H, W, image.stride(0), image.stride(1), sign, L.ptr(raw),
               raw_type(raw), raw.stride(0), raw.stride(1), L.ptr(coords[v]), L.ptr(counts[v:v + 1]), n[v],
               int(delta_fit), float(sigmaZ), float(sigmaXY), L.ptr(rows), st)
        out.append(rows)
    return out[0], out[1]
