"""Drift between acquisitions of one FOV (reference ``mermake/align.py``).

The heavy steps run in HIP: rasterisation of the point clouds at 1/resc scale,
the FFT cross-correlation (``mm_fft3d_c2c`` on a zero-padded smooth grid, which
leaves the linear correlation unchanged), a deterministic argmax (the
correlation of two binary images is an integer, so ``rint`` makes the peak
exact; ties go to the first index, as ``argmax`` does) and the five rounds of
nearest-neighbour refinement (brute force in float64, lowest index on distance
ties, replacing the cupyx KD-tree of ``align.py:8,32``).  The spectrum of the
reference raster stays resident in each ``Aligner`` (the reference recomputes it
in every ``fftconvolve`` call), so a FOV block's drifts share it.

Results are returned as NumPy arrays (``drift``'s records are pickled).
"""
import ctypes
import os
import pickle

import numpy as np
import torch

from . import _lib as L
from ._dev import as_device
from .utils import smooth_length


def _pts(X, device=None):
    X = as_device(X, device)
    if X.dtype != torch.float32:
        X = X.to(torch.float32)
    return X.contiguous()


class _Raster:
    """Rounded points and trim box of ``get_im_from_Xh`` (``align.py:37-59``)."""

    def __init__(self, X, resc, trim):
        n = X.shape[0]
        if n == 0:
            raise ValueError("zero-size array to reduction operation minimum which has no identity")
        dev = X.device
        self.Xi = torch.empty((n, 3), dtype=torch.int32, device=dev)
        b = torch.empty(6, dtype=torch.int32, device=dev)
        st = L.stream_ptr(dev)
        L.call("mm_points_round_bounds", L.ptr(X), n, X.stride(0), float(resc), L.ptr(self.Xi), L.ptr(b), st)
        bh = b.cpu().numpy()
        lo = bh[:3] + np.array([0, trim, trim])
        hi = bh[3:] - np.array([0, trim, trim])
        self.lo = (ctypes.c_int * 3)(*[int(v) for v in lo])
        self.hi = (ctypes.c_int * 3)(*[int(v) for v in hi])
        o4 = torch.empty(4, dtype=torch.int32, device=dev)
        L.call("mm_points_kept_max", L.ptr(self.Xi), n, self.lo, self.hi, L.ptr(o4), st)
        o = o4.cpu().numpy()
        self.n = n
        self.nkeep = int(o[3])
        if self.nkeep == 0:
            self.shape = (1, 1, 1)   # align.py:52-53: all-zero 1x1x1 image
        else:
            self.shape = tuple(int(v) + 1 for v in o[:3])
            # negative coordinates wrap as a NumPy index does (align.py:58); beyond -size it raises
            if any(s < 1 or int(lv) < -s for s, lv in zip(self.shape, bh[:3])):
                raise IndexError("rounded coordinate outside the image")

    def write(self, buf, flip=False):
        if self.nkeep == 0:
            return
        L0, L1, L2 = buf.shape
        L.call("mm_points_raster", L.ptr(self.Xi), self.n, self.lo, self.hi, L.ptr(buf), L0, L1, L2, int(flip),
               self.shape[0], self.shape[1], self.shape[2], L.stream_ptr(buf.device))


class Aligner:
    """``align.py:12-129``."""

    def __init__(self, X_ref, resc=5, trim=10, th=None):
        self.resc, self.trim, self.th = resc, trim, th
        X_ref = _pts(X_ref)
        if X_ref.ndim != 2 or X_ref.shape[1] < 3:
            raise IndexError("points need at least 3 columns")
        if th:
            X_ref = self.threshold(X_ref)
        if X_ref.numel() == 0:
            raise ValueError("No points left after thresholding; cannot build KDTree. "
                             "Check your data and/or threshold value.")
        self.X_ref = X_ref[:, :3].contiguous()
        self._ref = _Raster(self.X_ref, resc, trim)
        self.Xm_ref = np.zeros(3, dtype=np.int32)
        self._grid = None   # padded grid of the resident reference spectrum
        self._fb = None     # FFT of the flipped reference raster on that grid

    @property
    def im_ref(self):
        return self.get_im_from_Xh(self.X_ref)[0]

    def get_im_from_Xh(self, Xh):
        """Binary image and origin (``align.py:37-59``)."""
        r = _Raster(_pts(Xh)[:, :3].contiguous(), self.resc, self.trim)
        buf = torch.zeros(r.shape, dtype=torch.complex64, device=self.X_ref.device)
        r.write(buf)
        return buf.real.contiguous(), np.zeros(3, dtype=np.int32)

    def reference_spectrum(self, full):
        """FFT of the flipped reference raster on a padded grid of at least ``full`` per axis.

        The spectrum stays resident: it is computed once, on a grid with room for moving rasters a
        little larger than the reference's own (2 Rb + 2 per axis), and recomputed on a larger grid only
        when a correlation needs more.  A linear correlation computed on any grid of at least ``full`` is the
        same, so every correlation reuses it."""
        if self._grid is None or any(f > g for f, g in zip(full, self._grid)):
            Rb = self._ref.shape
            grid = tuple(smooth_length(max(f, 2 * b + 2)) for f, b in zip(full, Rb))
            fb = torch.zeros(grid, dtype=torch.complex64, device=self.X_ref.device)
            self._ref.write(fb, flip=True)
            L.call("mm_fft3d_c2c_box", L.ptr(fb), grid[0], grid[1], grid[2], Rb[0], Rb[1], L.stream_ptr(fb.device))
            self._grid, self._fb = grid, fb
        return self._grid, self._fb

    def correlation_peak(self, X):
        """Integer lag of the maximum of ``fftconvolve(im, im_ref[::-1,::-1,::-1])`` (``align.py:104-110``)."""
        r = _Raster(X, self.resc, self.trim)
        Ra, Rb = r.shape, self._ref.shape
        full = [a + b - 1 for a, b in zip(Ra, Rb)]
        grid, fb = self.reference_spectrum(full)
        dev = X.device
        st = L.stream_ptr(dev)
        fa = torch.zeros(grid, dtype=torch.complex64, device=dev)
        r.write(fa)
        L.call("mm_fft3d_c2c_box", L.ptr(fa), grid[0], grid[1], grid[2], Ra[0], Ra[1], st)   # the raster is in [0, Ra)
        L.call("mm_cmul", L.ptr(fa), L.ptr(fb), fa.numel(), st)
        L.call("mm_fft3d_c2c", L.ptr(fa), grid[0], grid[1], grid[2], 7, 1, st)
        ws = torch.empty(L.lib().mm_argmax_ws_bytes(), dtype=torch.uint8, device=dev)
        oi = torch.empty(1, dtype=torch.int64, device=dev)
        ov = torch.empty(1, dtype=torch.float32, device=dev)
        L.call("mm_argmax_rounded", L.ptr(fa), grid[0], grid[1], grid[2], full[0], full[1], full[2], L.ptr(oi),
               L.ptr(ov), L.ptr(ws), st)
        idx = np.array(np.unravel_index(int(oi.item()), full))
        return idx - np.array(Rb) + 1

    def get_Xtzxy(self, X, tzxy0, target=3):
        """Five nearest-neighbour refinement rounds (``align.py:61-77``)."""
        X = _pts(X)[:, :3].contiguous()
        dev = X.device
        n, m = X.shape[0], self.X_ref.shape[0]
        t = torch.tensor(np.asarray(tzxy0, dtype=np.float64), device=dev)
        ds = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        ix = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        npts = torch.zeros(1, dtype=torch.int32, device=dev)
        st = L.stream_ptr(dev)
        for dist_th in np.linspace(self.resc, target, 5):
            L.call("mm_nn_round", L.ptr(X), n, X.stride(0), L.ptr(self.X_ref), m, self.X_ref.stride(0), L.ptr(t),
                   float(dist_th), L.ptr(ds), L.ptr(ix), L.ptr(npts), st)
        return t.cpu().numpy(), int(npts.item())

    def get_best_translation_points(self, X, target=3, return_counts=False):
        """``align.py:94-118``."""
        X = _pts(X)
        if self.th:
            X = self.threshold(X)
        X = X[:, :3].contiguous()
        tzxy = self.correlation_peak(X) * self.resc
        tzxy, npts = self.get_Xtzxy(X, tzxy, target=target)
        return (tzxy, npts) if return_counts else tzxy

    def get_shifted_slices(self, X, shape):
        """``align.py:79-92``."""
        tzxy = np.round(self.get_best_translation_points(X)).astype(int).tolist()
        sig, bk = [], []
        for t, sh in zip(tzxy, shape):
            if t > 0:
                sig.append(slice(t, sh))
                bk.append(slice(0, sh - t))
            else:
                sig.append(slice(0, sh + t))
                bk.append(slice(-t, sh))
        return tuple(bk), tuple(sig)

    def threshold(self, X):
        """``align.py:121-129``: rows with ``X[:, -1] > th``."""
        X = _pts(X)
        dim = X.shape[1]
        if X.numel():
            X = X[X[:, -1] > self.th]
            if X.numel():
                return X
        return torch.zeros((0, dim), dtype=torch.float32, device=X.device)


class DualAligner:
    """``align.py:131-156``."""

    def __init__(self, ref, th=None):
        self.plus = Aligner(_pts(ref.Xh_plus)[:, :3], th=th)
        self.minus = Aligner(_pts(ref.Xh_minus)[:, :3], th=th)
        self.th = th

    def get_best_translation_pointsV2(self, obj):
        tp, n_p = self.plus.get_best_translation_points(obj.Xh_plus, return_counts=True)
        tm, n_m = self.minus.get_best_translation_points(obj.Xh_minus, return_counts=True)
        if n_p + n_m == 0:
            f = np.zeros(3)
        elif np.max(np.abs(tm - tp)) <= 2:
            f = -(tp * n_p + tm * n_m) / (n_p + n_m)
        else:
            f = -tp if n_p >= n_m else -tm
        return f, tp, tm, n_p, n_m


def drift_save(data, filepath):
    """``align.py:158-161``."""
    with open(filepath, "wb") as f:
        pickle.dump(data, f)


def drift(block):
    """``align.py:163-177``: ``([drifts, dirnames, ifov, ref.path], filepath)`` or None if the file exists."""
    ifov = block.ifov()
    block.iset()
    filepath = block.drift_file()
    if not os.path.exists(filepath):
        ref = block[len(block) // 2]
        dual = DualAligner(ref, th=4)
        drifts, files = [], []
        for image in block:
            drifts.append(dual.get_best_translation_pointsV2(image))
            files.append(os.path.dirname(image.path))
        return [drifts, files, ifov, ref.path], filepath
    return None
