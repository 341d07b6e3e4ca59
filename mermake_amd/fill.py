"""In-place single mirror reflection (reference ``mermake/fill.py:54-111``),
running ``mm_reflect``."""
import ctypes

import torch

from . import _lib as L
from ._dev import as_device


def reflect(arr, i, axis=0, mode="out", out=None):
    """``fill.reflect``: mode ``'out'`` mirrors positions after ``i`` from before it,
    ``'in'`` positions before ``i`` from after it; out-of-range mirrors are left as is."""
    arr = as_device(arr)
    if out is None:
        result = arr.clone()
    else:
        if tuple(out.shape) != tuple(arr.shape):
            raise ValueError(f"output array shape {tuple(out.shape)} does not match input shape {tuple(arr.shape)}")
        if out.dtype != arr.dtype:
            raise ValueError(f"output array dtype {out.dtype} does not match input dtype {arr.dtype}")
        result = out
        if result is not arr:
            result.copy_(arr)
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
        flag = 0
    elif mode == "in":
        flag = 1
    else:
        raise ValueError("mode must be 'out' or 'in'")
    if result.dtype != torch.float32:
        raise TypeError("reflect operates on float32 arrays")
    if result.ndim > 8:
        raise ValueError("at most 8 dimensions are supported")
    shape = (ctypes.c_longlong * result.ndim)(*result.shape)
    strides = (ctypes.c_longlong * result.ndim)(*result.stride())
    if result.numel():
        L.call("mm_reflect", L.ptr(result), result.ndim, shape, strides, int(axis), int(i), flag,
               L.stream_ptr(result.device))
    return result
