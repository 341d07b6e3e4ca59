"""Tiled Wiener deconvolution (reference ``mermake/deconvolver.py``), running the
``mm_deconv_*`` kernels.

``Deconvolver`` keeps the reference constructor and methods
(``deconvolver.py:47-236``).  Differences a caller can observe:

* ``psf_fft`` holds the half spectra ``(n_psf, Zp, T, T//2+1)`` (complex64),
  already divided by ``Zp*T*T`` (the ``ifftn`` normalisation), instead of the
  full ``(n_psf, Zp, T, T)`` spectra.  The PSF FFT behind it runs in float64.
  By default stacks of up to 48 planes hold the z-domain kernels ``ifft_z`` of
  those spectra followed by the spectra (with ``zpad = sz - 1``, sz <= 40: the
  240 planes of the length-80 z-pass operator; 40 planes with zpad 19 keep the
  spectra alone, for the length-78 z transforms), and tiles T = 113*M hold
  them column-blocked, ``(n_psf, planes, Hs/blk, T, blk)``
  (``mm_deconv_spec_layout``; ``set_schedule`` picks other forms).
* ``tile_wise`` computes tiles in batches; the yielded views stay valid until
  the generator computes the next batch (the reference overwrote them at every
  step).  ``tile_pad`` / ``tile_res`` are shape-only (``meta``) tensors.
* ``zpad < 1`` or ``overlap < 1`` raise ``ValueError`` in ``tile_wise``/``apply``
  (the reference produces empty slices there).
* a float64 flat field is rounded to float32 before the divide.
"""
import ctypes
import os
from itertools import cycle

import numpy as np
import torch

from . import _lib as L
from ._dev import as_device, default_device, unit_last_stride


def laplacian_3d(shape, device=None):
    """``deconvolver.py:24-36``: float16; 6 at the centre, -1 at z+-1, y+-1 and only x-1."""
    lap = torch.zeros(shape, dtype=torch.float16, device=device or default_device())
    zc, yc, xc = shape[0] // 2, shape[1] // 2, shape[2] // 2
    lap[zc, yc, xc] = 6
    lap[zc - 1, yc, xc] = -1
    lap[zc + 1, yc, xc] = -1
    lap[zc, yc - 1, xc] = -1
    lap[zc, yc + 1, xc] = -1
    lap[zc, yc, xc - 1] = -1
    return lap


def fftn_c2c(x, inverse=False):
    """3D complex64 FFT of a (n0, n1, n2) tensor through ``mm_fft3d_c2c`` (new tensor)."""
    y = x.to(torch.complex64).contiguous().clone()
    n0, n1, n2 = y.shape
    L.call("mm_fft3d_c2c", L.ptr(y), n0, n1, n2, 7, 1 if inverse else 0, L.stream_ptr(y.device))
    return y


def batch_laplacian_fft(batch_size, shape):
    """``deconvolver.py:38-44``: ``fftn(laplacian_3d(shape))[None]`` (complex64)."""
    return fftn_c2c(laplacian_3d(shape))[None]


# forms of the T-length passes and of the z pass (mm_deconv_set_schedule, in the order of MM_XY_* / MM_Z_*)
XY_FORMS = ("stockham", "prime_first", "rader")
Z_FORMS = ("fft", "matrix", "dft80")


class _Plan:
    def __init__(self, sz, tile_size, overlap, zpad):
        h = ctypes.c_void_p()
        L.call("mm_deconv_create", sz, tile_size, overlap, zpad, ctypes.byref(h))
        self.h = h
        T, Zp, H = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.call("mm_deconv_dims", h, ctypes.byref(T), ctypes.byref(Zp), ctypes.byref(H))
        self.T, self.Zp, self.H = T.value, Zp.value, H.value
        self.refresh()

    def refresh(self):
        """Read the spectrum layout of the current schedule."""
        h = self.h
        blk, Hs, wp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.call("mm_deconv_spec_layout", h, ctypes.byref(blk), ctypes.byref(Hs), ctypes.byref(wp))
        self.blk, self.Hs, self.wplanes = blk.value, Hs.value, wp.value
        self.spec_bytes = L.lib().mm_deconv_spec_bytes(h)

    def schedule(self):
        """``(xy form, z form, admitted xy forms, admitted z forms)`` as names (``XY_FORMS``, ``Z_FORMS``)."""
        v = [ctypes.c_int() for _ in range(4)]
        L.call("mm_deconv_schedule", self.h, *[ctypes.byref(x) for x in v])
        xy, z, xm, zm = (x.value for x in v)
        return (XY_FORMS[xy], Z_FORMS[z], [f for i, f in enumerate(XY_FORMS) if xm >> i & 1],
                [f for i, f in enumerate(Z_FORMS) if zm >> i & 1])

    def set_schedule(self, xy, z):
        L.call("mm_deconv_set_schedule", self.h, XY_FORMS.index(xy), Z_FORMS.index(z))
        self.refresh()

    def spectrum_shape(self, n):
        """Shape of ``n`` Wiener operators in the plan's layout: (n, planes, T, H) natural, or
        (n, planes, Hs // blk, T, blk) column-blocked (planes: Zp, or 240 for the length-80 z pass)."""
        if self.blk:
            return (n, self.wplanes, self.Hs // self.blk, self.T, self.blk)
        return (n, self.wplanes, self.T, self.H)

    def __del__(self):
        try:
            if self.h:
                L.lib().mm_deconv_destroy(self.h)
        except Exception:
            pass


def nearest_psf(cx, cy, psf_keys):
    """Index of the PSF whose key ``(..., x, y)`` is nearest to the point (cx, cy); the
    first such key on ties."""
    best, bd = 0, None
    for i, key in enumerate(psf_keys):
        d = (float(key[-2]) - cx) ** 2 + (float(key[-1]) - cy) ** 2
        if bd is None or d < bd:
            best, bd = i, d
    return best


def tile_geometry(shape, tile_size, overlap, n_psf=1, psf_keys=None, psf_mode="zip"):
    """Tiles of ``Deconvolver.tiled`` (``deconvolver.py:186-206``) with their padding
    parameters (``:117-144``).  With several PSFs, ``psf_mode`` pairs them with the tiles:
    ``"zip"`` as the reference (``deconvolver.py:110,116``: x-major order, truncated to the
    shorter list), ``"keyed"`` every tile with the PSF whose key ``(z, x, y)`` is nearest the
    centre of the tile's own region."""
    if psf_mode not in ("zip", "keyed"):
        raise ValueError(f"psf_mode must be 'zip' or 'keyed', not {psf_mode!r}")
    keyed = psf_mode == "keyed" and n_psf > 1
    if keyed and (psf_keys is None or len(psf_keys) != n_psf):
        raise ValueError("keyed PSF pairing needs one (z, x, y) key per PSF")
    sz, sx, sy = shape
    T = tile_size + 2 * overlap
    out = []
    k = 0
    for x in range(0, sx, tile_size):
        for y in range(0, sy, tile_size):
            if n_psf > 1 and not keyed and k >= n_psf:
                return out
            x0, x1 = max(x - overlap, 0), min(x + tile_size + overlap, sx)
            y0, y1 = max(y - overlap, 0), min(y + tile_size + overlap, sy)
            xd, yd = x1 - x0, y1 - y0
            xend = xd + (overlap if (x == 0 or xd < T) else 0)
            yend = yd + (overlap if (y == 0 or yd < T) else 0)
            out.append(dict(x=x, y=y, X0=x0, Y0=y0, xd=xd, yd=yd, xs=overlap if x == 0 else 0,
                            ys=overlap if y == 0 else 0, xfirst=int(x == 0), yfirst=int(y == 0),
                            xe=min(xend, T), ye=min(yend, T), psf=0 if n_psf == 1 else k))
            if keyed:
                out[-1]["psf"] = nearest_psf(x + (min(x + tile_size, sx) - x) / 2,
                                             y + (min(y + tile_size, sy) - y) / 2, psf_keys)
            k += 1
    return out


def tile_table(geom, image, flat):
    """Device array of ``mm_tile_desc`` for ``geom`` over ``image`` (and ``flat``)."""
    return torch.from_numpy(tile_records(geom, image, flat).view(np.uint8).copy()).to(image.device)


def tile_records(geom, image, flat):
    """Host array of ``mm_tile_desc`` records for ``geom`` over device tensors ``image`` / ``flat``."""
    arr = np.zeros(len(geom), dtype=L.TILE_DTYPE)
    for i, g in enumerate(geom):
        for k in ("X0", "Y0", "xd", "yd", "xs", "ys", "xfirst", "yfirst", "xe", "ye", "psf"):
            arr[i][k] = g[k]
        arr[i]["ox"], arr[i]["oy"] = g["x"], g["y"]
        arr[i]["img"] = image.data_ptr()
        arr[i]["img_sz"], arr[i]["img_sx"] = image.stride(0), image.stride(1)
        arr[i]["img_type"] = pixel_type(image)
        if flat is not None:
            arr[i]["flat"] = flat.data_ptr()
            arr[i]["flat_sx"] = flat.stride(0)
    return arr


class SquaredU8(torch.Tensor):
    """A uint8 stack of the microscope's raw values whose pixels are their squares, the
    ``astype(uint16) ** 2`` of ``read_im`` (io.py:140-142).  The tile loader squares them on the
    device, so such a stack is half the bytes of its uint16 form to copy and to read, with the same
    results.  Make one with ``SquaredU8.of(frames)``; ``.to(device)`` and views keep the type."""

    @staticmethod
    def of(frames):
        t = frames if isinstance(frames, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(frames))
        if t.dtype != torch.uint8:
            raise TypeError(f"SquaredU8 holds uint8 values, not {t.dtype}")
        return t.as_subclass(SquaredU8)


def pixel_type(image):
    """``mm_tile_desc.img_type`` of a prepared stack: 0 uint16, 1 float32, 2 squared uint8."""
    if isinstance(image, SquaredU8):
        return 2
    return 1 if image.dtype == torch.float32 else 0


def prepare_stack(image):
    """The stack as a device tensor the tile loader reads (uint16, float32 or ``SquaredU8``, unit
    y stride)."""
    image = as_device(image)
    if image.ndim != 3:
        raise ValueError("expected a (z, x, y) stack")
    if isinstance(image, SquaredU8):
        if image.dtype != torch.uint8:
            raise TypeError("a SquaredU8 stack must hold uint8 values")
        return unit_last_stride(image)
    if image.dtype not in (torch.uint16, torch.float32):
        image = image.to(torch.float32)   # tile_pad[...] = tile casts to float32 (deconvolver.py:124)
    return unit_last_stride(image)


def prepare_flat(flat_field, image):
    if flat_field is None:
        return None
    f = as_device(flat_field, image.device)
    if f.ndim == 3:
        if f.shape[0] != 1:
            raise NotImplementedError("per-plane flat fields are not supported; pass a 2D flat")
        f = f[0]
    if f.ndim != 2 or tuple(f.shape) != tuple(image.shape[1:]):
        raise ValueError(f"flat field shape {tuple(f.shape)} does not match the stack's {tuple(image.shape[1:])}")
    if f.dtype != torch.float32:
        f = f.to(torch.float32)
    return unit_last_stride(f)


class Deconvolver:
    """``mermake.deconvolver.Deconvolver`` on gfx950."""

    batch_bytes = int(os.environ.get("MERMAKE_BATCH_BYTES", str(24 << 30)))

    def __init__(self, psfs, channel_shape, tile_size=300, zpad=0, overlap=89, beta=0.001, xp=None, device=None,
                 psf_mode="zip", autotune=False, **kwargs):
        self.tile_size = tile_size
        self.psf_mode = psf_mode  # "zip" (the reference's pairing) or "keyed" (nearest PSF key per tile)
        self.tile_height = channel_shape[0]
        self.overlap = overlap
        self.zpad = zpad
        self.beta = beta
        self.device = torch.device(device) if device is not None else default_device()
        self.psf_keys = list(psfs.keys()) if isinstance(psfs, dict) else None
        if psf_mode not in ("zip", "keyed"):
            raise ValueError(f"psf_mode must be 'zip' or 'keyed', not {psf_mode!r}")
        self._psfs = psfs
        self.batch_tiles = None   # tiles per launch chosen by the autotuner (None: as many as memory allows)
        self.tuning = None
        self._plan = _Plan(self.tile_height, tile_size, overlap, zpad)
        self.psf_fft = None
        self.generation = 0       # counts the builds of psf_fft (captured kernel chains hold its address)
        self._ws = {}
        self._tables = {}         # tile tables of the stacks seen last, by address and layout
        self._build_spectra()
        if autotune:
            from .autotune import tune
            tune(self, channel_shape, **(autotune if isinstance(autotune, dict) else {}))

    def _operators(self, box_half=0):
        """The Wiener operators of every PSF in the layout of the plan's schedule; ``box_half > 0``
        folds the subtraction of the box blur of that half-width into them
        (``mm_deconv_build_wiener_box``)."""
        psfs = list(self._psfs.values()) if isinstance(self._psfs, dict) else [self._psfs]
        out = torch.empty(self._plan.spectrum_shape(len(psfs)), dtype=torch.complex64, device=self.device)
        ws_bytes = L.lib().mm_deconv_wiener_ws_bytes(self._plan.h)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.device)
        st = L.stream_ptr(self.device)
        for i, p in enumerate(psfs):
            c = self.center_psf(np.asarray(p))
            pd = torch.from_numpy(np.ascontiguousarray(c, dtype=np.float32)).to(self.device)
            L.call("mm_deconv_build_wiener_box", self._plan.h, L.ptr(pd), float(self.beta), int(box_half),
                   L.ptr(out[i]), L.ptr(ws), ws_bytes, st)
        torch.cuda.current_stream(self.device).synchronize()
        del ws
        return out

    def _build_spectra(self):
        """The Wiener operators of every PSF in the layout of the plan's schedule."""
        self.psf_fft = None   # the operators of the previous schedule go before the new ones come
        self._box_ops = {}
        self.psf_fft = self._operators()
        self.generation += 1

    def box_spectra(self, blur_radius):
        """``psf_fft`` with the blur subtraction of ``blur_radius`` (box half-width ``blur_radius // 2``)
        folded in, built on first use and kept with the operators of the current schedule."""
        half = int(blur_radius) // 2
        ops = self._box_ops.get(half)
        if ops is None:
            ops = self._box_ops[half] = self._operators(half)
        return ops

    def spectral_blur(self, blur_radius, window):
        """Whether the blur subtraction of ``blur_radius`` can ride on the Wiener operator for a
        result read only at rows and columns ``window = (lo, hi)``: every box around those voxels
        stays inside the tile (``lo >= half`` and ``hi + half <= T``), so the box on the periodic
        tile equals the reference's clipped box there (``blur.cu:16-45``)."""
        if blur_radius is None or window is None:
            return False
        half = int(blur_radius) // 2
        return half > 0 and window[0] >= half and window[1] + half <= self._plan.T

    def schedule(self):
        """The deconvolution schedule: ``{"xy": form, "z": form, "batch_tiles": n or None}`` and the
        forms the tile geometry admits (``"xy_forms"``, ``"z_forms"``); see ``set_schedule``."""
        xy, z, xs, zs = self._plan.schedule()
        return {"xy": xy, "z": z, "batch_tiles": self.batch_tiles, "xy_forms": xs, "z_forms": zs}

    def set_schedule(self, xy=None, z=None, batch_tiles=None):
        """Select the form of the T-length y/x FFT passes (``XY_FORMS``: Stockham radix passes,
        prime-first matrix stage + radix M, Rader prime stage) and of the z pass (``Z_FORMS``: z FFTs,
        folded z-kernel matrix, length-80 DFTs), and the tiles per launch.  Forms the geometry does
        not admit raise ``MermakeLibraryError``.  The Wiener operators are rebuilt in the new layout,
        and buffers sized by the old one (``HybServer`` slots) must be made again."""
        cur = self._plan.schedule()
        xy = cur[0] if xy is None else xy
        z = cur[1] if z is None else z
        if batch_tiles is not None:
            self.batch_tiles = max(1, int(batch_tiles))
        if (xy, z) != cur[:2]:
            if xy not in XY_FORMS or z not in Z_FORMS:
                raise ValueError(f"unknown forms ({xy!r}, {z!r}); xy in {XY_FORMS}, z in {Z_FORMS}")
            self._plan.set_schedule(xy, z)
            self._ws = {}
            self._build_spectra()

    def geometry(self, shape):
        """Tiles of a (z, x, y) channel shape with their PSF indices (``psf_mode``)."""
        return tile_geometry(tuple(shape), self.tile_size, self.overlap, self.psf_fft.shape[0], self.psf_keys,
                             self.psf_mode)

    # ------------------------------------------------------------ reference API
    @property
    def tile_pad(self):
        T = self._plan.T
        return torch.empty((2 * self.zpad + self.tile_height, T, T), dtype=torch.float32, device="meta")

    @property
    def tile_res(self):
        T = self._plan.T
        return torch.empty((self.tile_height, T, T), dtype=torch.float32, device="meta")

    def center_psf(self, psf):
        """``deconvolver.py:238-270`` (host NumPy, as in the reference)."""
        target = np.array([self.tile_height, self.tile_size, self.tile_size])
        shp = np.array(psf.shape)
        out = np.zeros(tuple(target.tolist()), dtype=psf.dtype)
        s_o = np.maximum(0, (target - shp) // 2)
        e_o = s_o + np.minimum(target, shp)
        s_p = np.maximum(0, (shp - target) // 2)
        e_p = s_p + np.minimum(target, shp)
        out[tuple(slice(int(a), int(b)) for a, b in zip(s_o, e_o))] = \
            psf[tuple(slice(int(a), int(b)) for a, b in zip(s_p, e_p))]
        out /= out.sum()
        return out

    def tiled(self, image):
        """``deconvolver.py:186-206``: yields ``(x, y, view)``."""
        sz, sx, sy = image.shape
        ts, ov = self.tile_size, self.overlap
        self.nx = -(-sx // ts)
        self.ny = -(-sy // ts)
        for x in range(0, sx, ts):
            for y in range(0, sy, ts):
                yield x, y, image[:, max(x - ov, 0):min(x + ts + ov, sx), max(y - ov, 0):min(y + ts + ov, sy)]

    def _check(self, image):
        if self.zpad < 1 or self.overlap < 1:
            raise ValueError("tile processing needs zpad >= 1 and overlap >= 1")
        if image.shape[0] != self.tile_height:
            raise ValueError(f"stack depth {image.shape[0]} does not match the Deconvolver's {self.tile_height}")

    def _tile_table(self, geom, img, flat):
        """``tile_table`` of a whole stack, cached by the stack's and flat's addresses and layout (a pageable
        host-to-device copy waits for the stream, so it is made once per stack, not per call).  The key holds
        the stream too: a table is only used on the stream it was allocated on, so dropping it is stream-ordered."""
        key = (img.data_ptr(), tuple(img.shape), tuple(img.stride()), img.dtype, type(img),
               None if flat is None else (flat.data_ptr(), tuple(flat.stride())), L.stream_ptr(self.device).value)
        t = self._tables.get(key)
        if t is None:
            if len(self._tables) >= 8:
                self._tables.clear()
            t = self._tables[key] = tile_table(geom, img, flat)
        return t

    def _workspace(self, name, nbytes):
        t = self._ws.get(name)
        if t is None or t.numel() < nbytes:
            t = None
            self._ws.pop(name, None)
            t = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._ws[name] = t
        return t

    def batch_size(self, ntiles, extra_per_tile=0):
        T = self._plan.T
        per = self._plan.spec_bytes + 4 * self.tile_height * T * T + extra_per_tile
        n = min(ntiles, self.batch_bytes // per)
        if self.batch_tiles:
            n = min(n, self.batch_tiles)
        return max(1, n)

    @staticmethod
    def fused_blur_maxima(blur_radius):
        """Whether ``run_tiles(..., maxima=...)`` can run the maxima inside the blur entry
        (the fused blur widths, box half-widths 15 and 25)."""
        return blur_radius is not None and int(blur_radius) // 2 in (15, 25)

    def run_tiles(self, table, ntiles, tiles_per_block, blur_radius, res_out=None, window=None, maxima=None,
                  spectral=False):
        """Deconvolution (+ blur subtraction) of ``ntiles`` descriptors: ``(ntiles, sz, T, T)``.

        ``spectral=True`` with a ``window`` that ``spectral_blur`` admits subtracts the blur through
        the operator (``box_spectra``): the deconvolution computes res - blur at rows and columns
        [lo, hi) directly, into ``res_out``, with no blur pass; ``maxima`` then rides on the
        deconvolution (``mm_tiles_deconv_maxima_window``: threshold bits from the inverse y transform).

        With a blur the result lives in the spectrum workspace (valid until the next call).
        ``window=(lo, hi)`` computes the blur subtraction only at rows and columns [lo, hi)
        (``mm_blur_sub_xy_window``), and the deconvolution only where those boxes read
        (``mm_deconv_tiles_window``); the rest of the result is left undefined.
        ``maxima=(threshold, delta, (lo, hi), coords, cap, counts, ws, ws_bytes)`` with a window
        and a fused blur width also finds the maxima of the result at rows and columns [lo, hi)
        in the same entry (``mm_tiles_blur_maxima_window``, the outputs of
        ``mm_tiles_maxima_window``)."""
        spectral = bool(spectral) and self.spectral_blur(blur_radius, window)
        if maxima is not None and (window is None or not (spectral or self.fused_blur_maxima(blur_radius))):
            raise ValueError("maxima in run_tiles need a window and the blur on the operator or of half-width "
                             "15 or 25")
        T, sz = self._plan.T, self.tile_height
        spec = self._workspace("spec", ntiles * self._plan.spec_bytes)
        res = res_out if res_out is not None else torch.empty((ntiles, sz, T, T), dtype=torch.float32,
                                                               device=self.device)
        st = L.stream_ptr(self.device)
        delta = None if blur_radius is None else int(blur_radius) // 2
        ops = self.box_spectra(blur_radius) if spectral else self.psf_fft
        args = (self._plan.h, L.ptr(table), ntiles, L.ptr(ops), ops.shape[0],
                int(tiles_per_block), L.ptr(spec), L.ptr(res))
        if spectral and maxima is not None:
            thr, mdelta, (lo, hi), coords, cap, counts, ws, ws_bytes = maxima
            L.call("mm_tiles_deconv_maxima_window", *args, int(window[0]), int(window[1]), float(thr), int(mdelta),
                   int(lo), int(hi), L.ptr(coords), int(cap), L.ptr(counts), L.ptr(ws), int(ws_bytes), st)
            return res
        if spectral:
            L.call("mm_deconv_tiles_window", *args, int(window[0]), int(window[1]), st)
            return res
        if window is None:
            L.call("mm_deconv_tiles", *args, st)
        else:
            d = 0 if delta is None else delta
            L.call("mm_deconv_tiles_window", *args, max(0, window[0] - d), min(T, window[1] + d), st)
        if blur_radius is None:
            return res
        # deconvolver.py:160-163: res - box_y(box_x(res)); the spectrum buffer is free now
        out = spec[:ntiles * sz * T * T * 4].view(torch.float32).view(ntiles, sz, T, T)
        tmp = None
        if delta not in (15, 25):
            tmp = self._workspace("blur_tmp", ntiles * sz * T * T * 4)
        if window is None:
            L.call("mm_blur_sub_xy", L.ptr(res), L.ptr(out), L.ptr(tmp), ntiles * sz, T, delta, st)
        elif maxima is not None:
            thr, mdelta, (lo, hi), coords, cap, counts, ws, ws_bytes = maxima
            L.call("mm_tiles_blur_maxima_window", self._plan.h, L.ptr(table), ntiles, L.ptr(res), L.ptr(out), delta,
                   int(window[0]), int(window[1]), float(thr), int(mdelta), int(lo), int(hi), L.ptr(coords), int(cap),
                   L.ptr(counts), L.ptr(ws), int(ws_bytes), st)
        else:
            L.call("mm_blur_sub_xy_window", L.ptr(res), L.ptr(out), L.ptr(tmp), ntiles * sz, T, delta,
                   int(window[0]), int(window[1]), st)
        return out

    def _batches(self, image, flat_field, blur_radius, want_raw, window=None, spectral=False):
        self._check(image)
        img = prepare_stack(image)
        flat = prepare_flat(flat_field, img)
        geom = self.geometry(img.shape)
        T, sz = self._plan.T, self.tile_height
        single = self.psf_fft.shape[0] == 1
        B = self.batch_size(len(geom), 4 * sz * T * T if want_raw else 0)
        table_all = self._tile_table(geom, img, flat)
        rec = L.TILE_DTYPE.itemsize
        for b0 in range(0, len(geom), B):
            g = geom[b0:b0 + B]
            table = table_all[b0 * rec:(b0 + len(g)) * rec]
            res = self.run_tiles(table, len(g), len(g) if single else 1, blur_radius, window=window,
                                 spectral=spectral)
            raw = None
            if want_raw:
                raw = torch.empty((len(g), sz, T, T), dtype=torch.float32, device=self.device)
                L.call("mm_deconv_tile_raw", self._plan.h, L.ptr(table), len(g), L.ptr(raw),
                       L.stream_ptr(self.device))
            yield g, table, res, raw, img

    def tile_wise(self, image, flat_field=None, blur_radius=None, **kwargs):
        """``deconvolver.py:98-170``: yields ``(x, y, tile, raw)`` views of shape (sz, xend, yend)."""
        for g, _, res, raw, _ in self._batches(image, flat_field, blur_radius, True):
            for i, t in enumerate(g):
                yield t["x"], t["y"], res[i, :, :t["xe"], :t["ye"]], raw[i, :, :t["xe"], :t["ye"]]

    def apply(self, image, flat_field=None, blur_radius=None, output=None, spectral=False, **kwargs):
        """``deconvolver.py:172-184``: stitched float32 volume.  ``spectral=True`` subtracts the blur
        through the operator where ``spectral_blur`` admits the stitched window (the DAPI caller)."""
        if output is None:
            output = torch.empty(tuple(image.shape), dtype=torch.float32, device=self.device)
        if output.dtype != torch.float32 or output.stride(-1) != 1:
            raise ValueError("output must be float32 with unit stride along y")
        sz, sx, sy = image.shape
        # the stitch reads rows and columns [ov, T - ov) of every tile: only those are computed
        T, ov = self._plan.T, self.overlap
        spectral = bool(spectral) and self.spectral_blur(blur_radius, (ov, T - ov))
        if self._plan.schedule()[0] == "rader" and (blur_radius is None or spectral):
            # the inverse y transform stores the owned parts straight into the output
            self._check(image)
            img = prepare_stack(image)
            flat = prepare_flat(flat_field, img)
            geom = self.geometry(img.shape)
            ops = self.box_spectra(blur_radius) if spectral else self.psf_fft
            single = ops.shape[0] == 1
            B = self.batch_size(len(geom))
            table_all = self._tile_table(geom, img, flat)
            rec = L.TILE_DTYPE.itemsize
            for b0 in range(0, len(geom), B):
                g = geom[b0:b0 + B]
                table = table_all[b0 * rec:(b0 + len(g)) * rec]
                spec = self._workspace("spec", len(g) * self._plan.spec_bytes)
                L.call("mm_deconv_tiles_stitch", self._plan.h, L.ptr(table), len(g), L.ptr(ops), ops.shape[0],
                       len(g) if single else 1, L.ptr(spec), L.ptr(output), output.stride(0), output.stride(1), sx,
                       sy, L.stream_ptr(self.device))
            return output
        for g, table, res, _, _ in self._batches(image, flat_field, blur_radius, False, window=(ov, T - ov),
                                                 spectral=spectral):
            L.call("mm_deconv_stitch", self._plan.h, L.ptr(table), len(g), L.ptr(res), L.ptr(output),
                   output.stride(0), output.stride(1), sx, sy, L.stream_ptr(self.device))
        return output


def full_deconv(image, psfs, flat_field=None, tile_size=300, zpad=None, overlap=1, beta=0.001):
    """``deconvolver.py:272-287``."""
    shape = tuple(image.shape)
    if zpad is None:
        zpad = shape[0]
    d = Deconvolver(psfs, channel_shape=shape, zpad=zpad, tile_size=tile_size, overlap=overlap, beta=beta)
    out = d.apply(image, flat_field)
    del d
    return out
