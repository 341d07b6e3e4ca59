"""FFT convolution (reference ``mermake/utils.py:10-36``) on ``mm_fft3d_c2c``.

The linear convolution is computed on a zero-padded grid whose lengths are the
next 2^a 3^b 5^c 7^d >= ``sa + sb - 1`` (a linear convolution does not depend
on how far it is padded), then cropped to the 'full' size.  Inputs are real, of
any rank (as the reference's ``fftn(a, shape)``); the result is float32.
"""
import torch

from . import _lib as L
from ._dev import as_device


def smooth_length(n):
    """Smallest 2^a 3^b 5^c 7^d >= n."""
    m = max(1, int(n))
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def _fftn(x, inverse=False):
    """In-place complex64 FFT over every axis of ``x`` (any rank >= 1): one ``mm_fft3d_c2c`` call over
    all three axes of a 3D array, otherwise one call per axis on the (outer, n, inner) view of the axis."""
    st = L.stream_ptr(x.device)
    if x.ndim == 3:
        n0, n1, n2 = x.shape
        L.call("mm_fft3d_c2c", L.ptr(x), n0, n1, n2, 7, 1 if inverse else 0, st)
        return x
    shape = list(x.shape)
    for k, n in enumerate(shape):
        if n == 1:
            continue
        outer = 1
        for v in shape[:k]:
            outer *= v
        inner = 1
        for v in shape[k + 1:]:
            inner *= v
        L.call("mm_fft3d_c2c", L.ptr(x), outer, n, inner, 2, 1 if inverse else 0, st)
    return x


def fftconvolve(a, b, mode="full"):
    """``utils.fftconvolve`` for real arrays of any rank (``fftn(a, shape)`` over every axis):
    'full', 'same' or 'valid'."""
    a = as_device(a)
    b = as_device(b, a.device)
    if a.ndim != b.ndim or a.ndim < 1:
        raise ValueError("fftconvolve expects two arrays of the same rank (at least 1)")
    if mode not in ("full", "same", "valid"):
        raise ValueError("mode must be 'full', 'same', or 'valid'")
    nd = a.ndim
    full = [sa + sb - 1 for sa, sb in zip(a.shape, b.shape)]
    grid = [smooth_length(f) for f in full]
    fa = torch.zeros(grid, dtype=torch.complex64, device=a.device)
    fb = torch.zeros(grid, dtype=torch.complex64, device=a.device)
    fa.real[tuple(slice(0, n) for n in a.shape)] = a.to(torch.float32)
    fb.real[tuple(slice(0, n) for n in b.shape)] = b.to(torch.float32)
    _fftn(fa)
    _fftn(fb)
    L.call("mm_cmul", L.ptr(fa), L.ptr(fb), fa.numel(), L.stream_ptr(a.device))
    _fftn(fa, inverse=True)
    out = fa.real[tuple(slice(0, f) for f in full)].contiguous()
    if mode == "full":
        return out
    if mode == "same":
        st = [(o - s) // 2 for o, s in zip(out.shape, a.shape)]
        return out[tuple(slice(st[i], st[i] + a.shape[i]) for i in range(nd))]
    vs = [sa - sb + 1 for sa, sb in zip(a.shape, b.shape)]
    st = [sb - 1 for sb in b.shape]
    return out[tuple(slice(st[i], st[i] + vs[i]) for i in range(nd))]
