"""FFT convolution (reference ``mermake/utils.py:10-36``) on ``mm_fft3d_c2c``.

The linear convolution is computed on a zero-padded grid whose lengths are the
next 2^a 3^b 5^c 7^d >= ``sa + sb - 1`` (a linear convolution does not depend
on how far it is padded), then cropped to the 'full' size.  Inputs are real;
the result is float32.
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


def _fft3(x, inverse=False):
    n0, n1, n2 = x.shape
    L.call("mm_fft3d_c2c", L.ptr(x), n0, n1, n2, 7, 1 if inverse else 0, L.stream_ptr(x.device))
    return x


def fftconvolve(a, b, mode="full"):
    """``utils.fftconvolve`` for 3D real arrays: 'full', 'same' or 'valid'."""
    a = as_device(a)
    b = as_device(b, a.device)
    if a.ndim != 3 or b.ndim != 3:
        raise ValueError("fftconvolve expects 3D arrays")
    full = [sa + sb - 1 for sa, sb in zip(a.shape, b.shape)]
    grid = [smooth_length(f) for f in full]
    fa = torch.zeros(grid, dtype=torch.complex64, device=a.device)
    fb = torch.zeros(grid, dtype=torch.complex64, device=a.device)
    fa.real[:a.shape[0], :a.shape[1], :a.shape[2]] = a.to(torch.float32)
    fb.real[:b.shape[0], :b.shape[1], :b.shape[2]] = b.to(torch.float32)
    _fft3(fa)
    _fft3(fb)
    L.call("mm_cmul", L.ptr(fa), L.ptr(fb), fa.numel(), L.stream_ptr(a.device))
    _fft3(fa, inverse=True)
    out = fa.real[:full[0], :full[1], :full[2]].contiguous()
    if mode == "full":
        return out
    if mode == "same":
        st = [(o - s) // 2 for o, s in zip(out.shape, a.shape)]
        return out[tuple(slice(st[i], st[i] + a.shape[i]) for i in range(3))]
    if mode == "valid":
        vs = [sa - sb + 1 for sa, sb in zip(a.shape, b.shape)]
        st = [sb - 1 for sb in b.shape]
        return out[tuple(slice(st[i], st[i] + vs[i]) for i in range(3))]
    raise ValueError("mode must be 'full', 'same', or 'valid'")
