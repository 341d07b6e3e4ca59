"""Oracle: the host-feed arithmetic of ``mermake/io.py`` (TEST INFRASTRUCTURE ONLY).

* ``channels_from_frames`` restates ``read_im`` (``io.py:116-146``) on already decoded
  frames: frame 0 dropped, the first ``(n // nch) * nch`` frames reshaped to
  ``(n // nch, nch, H, W)`` and swapped to channel-major, uint8 squared into uint16.
* ``load_flat`` restates one flat of ``load_flats`` (``io.py:42-54``): centre crop
  (``io.py:34-40``), ``box((20, 20))`` over the last two axes (``oracle.blur.box``), division
  by the median (``numpy.median``: the middle value, or the float32 mean of the two middle
  values; ``cupy.median`` computes the same).

Pins: ``channels_from_frames`` against the input windows of ``tests/golden/fixture_tiles.npz``
(decoded from the reference's fixture zarr and validated through the deconvolution golden
test) and the channel-major index identity ``out[c, k] = frames[1 + k * nch + c]``;
``load_flat`` through the pinned ``oracle.blur.box``, the closed form of a constant flat (all
ones) and a brute-force median by full sort.
"""
import numpy as np

from . import blur as _blur


def channels_from_frames(frames, nch):
    frames = np.asarray(frames)[1:]
    n = frames.shape[0]
    nzs = (n // nch) * nch
    im = frames[:nzs].reshape(n // nch, nch, frames.shape[-2], frames.shape[-1]).swapaxes(0, 1)
    if im.dtype == np.uint8:
        im = im.astype(np.uint16) ** 2
    return im


def center_crop(A, shape):
    h, w = shape[-2:]
    H, W = A.shape[-2:]
    top, left = (H - h) // 2, (W - w) // 2
    return A[top:top + h, left:left + w]


def load_flat(im, shape=None):
    if shape is not None:
        im = center_crop(im, shape)
    cim = np.ascontiguousarray(im, dtype=np.float32)
    blurred = _blur.box(cim, (20, 20), axes=(-1, -2))
    med = np.float32(np.median(blurred))
    return (blurred / med).astype(np.float32)
