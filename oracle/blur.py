"""Oracle: truncated-window box mean along one axis (TEST INFRASTRUCTURE ONLY).

Follows ``mermake/blur.py:15-188`` and the CUDA kernel ``mermake/blur.cu:1-52``.

Exact semantics: ``delta = size // 2``; each output is the mean of the input over
``[i - delta, i + delta]`` clipped to the axis, accumulated **sequentially in
float32 from the low to the high index**, then ``sum / num`` (one IEEE division).
The vectorised form below adds the shifted arrays in exactly that order, so the
float32 result is bit-identical to the sequential loop.

Pins: the exact 3x3x3 values of ``tests/test_blur.py:62-100`` and the NumPy
reference ``numpy_box_1d`` of ``tests/test_blur.py:9-37`` (replayed in
``tests/test_oracle_cpu.py``), plus a brute-force per-element loop on random
arrays.
"""
import numpy as np


def box_1d(image, size, axis=0, out=None, dtype=np.float32):
    """Restatement of ``blur.box_1d`` (``blur.py:157-188``)."""
    image = np.asarray(image)
    if image.dtype != dtype:
        raise AssertionError("Input must be float32" if dtype == np.float32 else "dtype mismatch")
    if image.ndim not in (2, 3):
        raise ValueError("Only 2D or 3D arrays are supported")
    if axis < 0:
        axis = image.ndim - 1          # blur.py:167-168: any negative axis -> last axis
    if image.ndim == 2 and axis >= 2:
        raise AssertionError(f"axis {axis} is out of bounds for array of dimension 2")
    delta = size // 2
    n = image.shape[axis]
    x = np.moveaxis(image, axis, 0)
    acc = np.zeros_like(x)
    num = np.zeros(n, dtype=np.int64)
    for k in range(-delta, delta + 1):     # window offsets low to high: the sequential sum order
        lo, hi = max(0, -k), min(n, n - k)  # outputs i whose input i + k lies on the axis
        if lo < hi:
            acc[lo:hi] += x[lo + k:hi + k]
            num[lo:hi] += 1
    shape = [n] + [1] * (x.ndim - 1)
    res = (acc / num.reshape(shape).astype(dtype)).astype(dtype)
    res = np.moveaxis(res, 0, axis)
    if out is not None:
        out[...] = res
        return out
    return np.ascontiguousarray(res)


def box(image, size, axes=None, out=None, temp=None):
    """Restatement of ``blur.box`` (``blur.py:15-129``): sorted, de-duplicated axes, applied in order."""
    image = np.asarray(image)
    if image.dtype != np.float32:
        image = image.astype(np.float32)
    if image.ndim not in (2, 3):
        raise ValueError("Only 2D or 3D arrays are supported")
    if axes is None:
        axes = tuple(range(image.ndim))
    elif isinstance(axes, int):
        axes = (axes,)
    elif not isinstance(axes, (tuple, list)):
        raise TypeError("axes must be None, int, tuple, or list")
    norm = []
    for a in axes:
        if a < 0:
            a = image.ndim + a
        if a < 0 or a >= image.ndim:
            raise ValueError(f"axis {a} is out of bounds for array of dimension {image.ndim}")
        norm.append(a)
    axes = tuple(sorted(set(norm)))
    if len(axes) == 0:
        return image.copy() if out is None else np.copyto(out, image) or out
    if isinstance(size, int):
        sizes = {a: size for a in axes}
    elif isinstance(size, (tuple, list)):
        if len(size) != len(axes):
            raise ValueError(f"Size must have {len(axes)} elements for {len(axes)} axes")
        sizes = dict(zip(axes, size))
    else:
        raise TypeError("size must be an int, tuple, or list of ints")
    res = image
    for a in axes:
        res = box_1d(res, sizes[a], axis=a)
    if out is not None:
        out[...] = res
        return out
    return res


def box_2d(image, size, axes=(-2, -1), out=None):
    """Restatement of ``blur.box_2d`` (``blur.py:132-155``)."""
    if len(axes) != 2:
        raise ValueError("axes must contain exactly 2 elements for 2D blur")
    return box(image, size, axes=axes, out=out)


def box_1d_bruteforce(arr, size, axis=0):
    """Per-element sequential float32 loop (the CUDA kernel's order), for pinning ``box_1d``."""
    arr = np.asarray(arr, dtype=np.float32)
    out = np.empty_like(arr)
    delta = size // 2
    n = arr.shape[axis]
    for idx in np.ndindex(arr.shape):
        i = idx[axis]
        s = np.float32(0.0)
        num = 0
        for j in range(max(0, i - delta), min(n - 1, i + delta) + 1):
            jj = list(idx)
            jj[axis] = j
            s = np.float32(s + arr[tuple(jj)])
            num += 1
        out[idx] = np.float32(s / np.float32(num))
    return out
