"""Separable box blur (reference ``mermake/blur.py``), running ``mm_box_1d``.

Signatures, axis/size handling and exceptions follow ``mermake/blur.py:15-188``;
the arithmetic is the reference kernel's (``mermake/blur.cu:1-52``): sequential
float32 window sums divided by the element count.
"""
import torch

from . import _lib as L
from ._dev import as_device, shares_memory


def box_1d(image, size, axis=0, out=None):
    """``blur.box_1d`` (``blur.py:157-188``)."""
    image = as_device(image)
    assert image.dtype == torch.float32, "Input must be float32"
    if out is None:
        out = torch.empty(image.shape, dtype=torch.float32, device=image.device)
    assert out.dtype == torch.float32, "Output must be float32"
    if shares_memory(image, out):
        raise ValueError("In-place operation not supported - input and output must be different")
    if axis < 0:
        axis = image.ndim - 1          # blur.py:167-168: any negative axis maps to the last one
    delta = size // 2
    if image.ndim == 2:
        sx, sy = image.shape
        sz = 1
        assert axis < 2, f"axis {axis} is out of bounds for array of dimension 2"
        axis += 1
    elif image.ndim == 3:
        sz, sx, sy = image.shape
    else:
        raise ValueError("Only 2D or 3D arrays are supported")
    src = image.contiguous()
    dst = out if out.is_contiguous() else torch.empty(out.shape, dtype=torch.float32, device=out.device)
    L.call("mm_box_1d", L.ptr(src), L.ptr(dst), sz, sx, sy, int(delta), int(axis), L.stream_ptr(image.device))
    if dst is not out:
        out.copy_(dst)
    return out


def box(image, size, axes=None, out=None, temp=None):
    """``blur.box`` (``blur.py:15-129``): axes sorted and de-duplicated, ping-pong through ``temp``."""
    image = as_device(image)
    if image.dtype != torch.float32:
        image = image.to(torch.float32)
    if image.ndim not in (2, 3):
        raise ValueError("Only 2D or 3D arrays are supported")
    if axes is None:
        axes = tuple(range(image.ndim))
    elif isinstance(axes, int):
        axes = (axes,)
    elif not isinstance(axes, (tuple, list)):
        raise TypeError("axes must be None, int, tuple, or list")
    norm = []
    for axis in axes:
        if axis < 0:
            axis = image.ndim + axis
        if axis < 0 or axis >= image.ndim:
            raise ValueError(f"axis {axis} is out of bounds for array of dimension {image.ndim}")
        norm.append(axis)
    axes = tuple(sorted(set(norm)))
    if len(axes) == 0:
        if out is None:
            return image.clone()
        out.copy_(image)
        return out
    if isinstance(size, int):
        sizes = {a: size for a in axes}
    elif isinstance(size, (tuple, list)):
        if len(size) != len(axes):
            raise ValueError(f"Size must have {len(axes)} elements for {len(axes)} axes")
        sizes = dict(zip(axes, size))
    else:
        raise TypeError("size must be an int, tuple, or list of ints")
    if out is None:
        out = torch.empty_like(image)
    if len(axes) == 1:
        box_1d(image, sizes[axes[0]], axis=axes[0], out=out)
        return out
    if temp is None:
        temp = torch.empty_like(image)
    elif temp.shape != image.shape or temp.dtype != image.dtype:
        raise ValueError("temp buffer must have same shape and dtype as input")
    box_1d(image, sizes[axes[0]], axis=axes[0], out=out)
    for i in range(1, len(axes)):
        a = axes[i]
        if i % 2 == 1:
            box_1d(out, sizes[a], axis=a, out=temp)
        else:
            box_1d(temp, sizes[a], axis=a, out=out)
    if len(axes) % 2 == 0:
        out.copy_(temp)
    return out


def box_2d(image, size, axes=(-2, -1), out=None):
    """``blur.box_2d`` (``blur.py:132-155``)."""
    if len(axes) != 2:
        raise ValueError("axes must contain exactly 2 elements for 2D blur")
    return box(image, size, axes=axes, out=out)
