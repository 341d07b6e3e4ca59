"""Oracle: in-place single mirror reflection about an index (TEST INFRASTRUCTURE ONLY).

Follows ``mermake/fill.py:6-111`` (CUDA ``mirror`` kernel + ``reflect`` wrapper).

Semantics (``fill.py:21-40``): for mode ``'in'`` every position ``j < i`` takes the
value at ``2i - j`` when that index is inside the axis; for mode ``'out'`` every
position ``j > i`` takes the value at ``2i - j`` when that index is ``>= 0``.
Positions whose mirror falls outside the axis are left unchanged.  Each mode
reads only one side of ``i`` and writes only the other, so the in-place update
is order independent.

Pins: the exact arrays of the reference tests ``tests/test_fill.py:22-133``
(including the ``i=3, mode='in'`` edge case) and the validation errors of
``tests/test_fill.py:135-177`` are replayed in ``tests/test_oracle_cpu.py``.
"""
import numpy as np


def reflect(arr, i, axis=0, mode="out", out=None):
    """NumPy restatement of ``mermake.fill.reflect`` (``fill.py:54-111``)."""
    if out is None:
        result = np.array(arr, copy=True)
    else:
        if out.shape != arr.shape:
            raise ValueError(f"output array shape {out.shape} does not match input shape {arr.shape}")
        if out.dtype != arr.dtype:
            raise ValueError(f"output array dtype {out.dtype} does not match input dtype {arr.dtype}")
        result = out
        if result is not arr:
            np.copyto(result, arr)
    if axis < 0:
        axis += result.ndim
    if axis < 0 or axis >= result.ndim:
        raise IndexError(f"axis {axis} is out of bounds for array of dimension {result.ndim}")
    n = result.shape[axis]
    if i < 0:
        i += n
    if i < 0 or i >= n:
        raise IndexError(f"index {i} is out of bounds for axis {axis} with size {n}")
    if mode == "out":
        dst = np.arange(i + 1, n)
        src = 2 * i - dst
        ok = src >= 0
    elif mode == "in":
        dst = np.arange(0, i)
        src = 2 * i - dst
        ok = src < n
    else:
        raise ValueError("mode must be 'out' or 'in'")
    dst, src = dst[ok], src[ok]
    if dst.size:
        moved = np.moveaxis(result, axis, 0)
        moved[dst] = moved[src]
    return result


def reflect_map(n, i, mode):
    """Index map of one reflection: ``m[j]`` = source index for position ``j`` (itself if untouched)."""
    m = np.arange(n)
    if mode == "out":
        j = np.arange(i + 1, n)
        s = 2 * i - j
        ok = s >= 0
    else:
        j = np.arange(0, i)
        s = 2 * i - j
        ok = s < n
    m[j[ok]] = s[ok]
    return m
