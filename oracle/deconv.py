"""Oracle: tiled Wiener deconvolution + hyb/DAPI callers (TEST INFRASTRUCTURE ONLY).

Follows ``mermake/deconvolver.py:24-44`` (Laplacian), ``:47-96`` (Wiener spectra),
``:98-170`` (``tile_wise``), ``:172-184`` (``apply``), ``:186-206`` (``tiled``),
``:238-270`` (``center_psf``), ``:272-287`` (``full_deconv``), and the callers
``mermake/__main__.py:184,195-198`` (zpad), ``:213-229`` (DAPI) and
``:242-269`` (hyb tile loop, keep-filter, offsets).

Modes: ``"fp64"`` carries every step in float64/complex128 (accuracy reference);
``"fp32"`` mirrors the reference dtypes (float32 tile, complex64 spectra, float16
Laplacian) and is the CPU throughput baseline.

Padding is emulated with the same *sequential in-place* reflections as the
reference (x, then y, then z-in, then z-out), so the zero planes that appear when
``zpad >= sz`` are reproduced.  ``padding_maps`` gives the equivalent closed-form
per-axis source-index maps used by the HIP tile loader; a test checks the two
against each other on every tile class.

Pins: ``tests/examples/single/out/deconv.300.*.npz`` (the reference's golden
float16 deconvolution of the real fixture, ``tests/test_deconvolver.py:422-434``)
through committed tile windows under ``tests/golden/``; closed forms for the
Wiener spectrum (delta PSF -> ``W = 1/(1+beta|L|^2)``, ``L(0) = 1``);
``tests/test_deconvolver.py:158-218`` (``center_psf`` and ``tiled``); the flat field
by linearity (deconvolving v with flat f equals deconvolving v/f without one, with or
without blur; the raw views hold (v/f)*f on the data mask and v/f in the pads).
"""
import os
from itertools import cycle

import numpy as np

from . import blur as oblur
from . import fill as ofill
from . import maxima as omax

try:
    import scipy.fft as _sfft
    _WORKERS = len(os.sched_getaffinity(0))

    def _fftn(a):
        return _sfft.fftn(a, workers=_WORKERS)

    def _ifftn(a):
        return _sfft.ifftn(a, workers=_WORKERS)
except Exception:  # pragma: no cover
    _fftn, _ifftn = np.fft.fftn, np.fft.ifftn


def laplacian_3d(shape):
    """``deconvolver.py:24-36``: float16, 6 at the centre, -1 at z+-1, y+-1 and only x-1 (sums to 1)."""
    lap = np.zeros(shape, dtype=np.float16)
    zc, yc, xc = shape[0] // 2, shape[1] // 2, shape[2] // 2
    lap[zc, yc, xc] = 6
    lap[zc - 1, yc, xc] = -1
    lap[zc + 1, yc, xc] = -1
    lap[zc, yc - 1, xc] = -1
    lap[zc, yc + 1, xc] = -1
    lap[zc, yc, xc - 1] = -1
    lap[zc, yc, xc - 1] = -1
    return lap


def laplacian_power(shape):
    """Closed form of ``|fftn(laplacian_3d(shape))|^2`` in float64 (position phase cancels)."""
    k0 = 2 * np.pi * np.arange(shape[0]) / shape[0]
    k1 = 2 * np.pi * np.arange(shape[1]) / shape[1]
    k2 = 2 * np.pi * np.arange(shape[2]) / shape[2]
    re = 6 - 2 * np.cos(k0)[:, None, None] - 2 * np.cos(k1)[None, :, None] - np.cos(k2)[None, None, :]
    im = np.broadcast_to(np.sin(k2)[None, None, :], re.shape)
    return re * re + im * im


def center_psf(psf, tile_height, tile_size):
    """``deconvolver.py:238-270``: centre/crop into (sz, ts, ts), normalise by the sum in the PSF dtype."""
    target = np.array([tile_height, tile_size, tile_size])
    shp = np.array(psf.shape)
    out = np.zeros(tuple(target.tolist()), dtype=psf.dtype)
    s_o = np.maximum(0, (target - shp) // 2)
    e_o = s_o + np.minimum(target, shp)
    s_p = np.maximum(0, (shp - target) // 2)
    e_p = s_p + np.minimum(target, shp)
    out[tuple(slice(int(a), int(b)) for a, b in zip(s_o, e_o))] = psf[tuple(slice(int(a), int(b)) for a, b in zip(s_p, e_p))]
    out /= out.sum()
    return out


def psf_stack(psfs, sz, tile_size, zpad, overlap):
    """``deconvolver.py:52-62``: centred PSFs, zero-padded, rolled by ``(-shape)//2``."""
    if isinstance(psfs, dict):
        st = np.stack([center_psf(p, sz, tile_size) for p in psfs.values()])
    else:
        st = center_psf(psfs, sz, tile_size)[None, ...]
    st = np.pad(st, ((0, 0), (zpad, zpad), (overlap, overlap), (overlap, overlap)), mode="constant")
    shift = -np.array(st.shape[1:]) // 2
    return np.roll(st, shift=tuple(shift.tolist()), axis=(1, 2, 3))


def wiener(psfs, sz, tile_size, zpad, overlap, beta, mode="fp64"):
    """Wiener spectra ``W = conj(H) / (H conj(H) + beta |L|^2)`` (``deconvolver.py:64-74``), one per PSF."""
    st = psf_stack(psfs, sz, tile_size, zpad, overlap)
    out = []
    for p in st:
        if mode == "fp64":
            H = _fftn(p.astype(np.float64))
            L2 = laplacian_power(p.shape)
            out.append(np.conj(H) / (H * np.conj(H) + beta * L2))
        else:
            H = _fftn(p.astype(np.float32)).astype(np.complex64)
            Hc = np.conj(H)
            H *= Hc
            L = _fftn(laplacian_3d(p.shape).astype(np.float32)).astype(np.complex64)
            L *= np.conj(L)
            L *= beta
            H += L
            out.append((Hc / H).astype(np.complex64))
    return out


def tiled(shape, tile_size, overlap):
    """``deconvolver.py:186-206``: x-major tile origins with their overlapped windows."""
    sz, sx, sy = shape
    for x in range(0, sx, tile_size):
        for y in range(0, sy, tile_size):
            yield (x, y, max(x - overlap, 0), min(x + tile_size + overlap, sx),
                   max(y - overlap, 0), min(y + tile_size + overlap, sy))


def padding_maps(T, start, d, first):
    """Closed form of one axis' reflection (``deconvolver.py:131-144``): padded index -> data index or -1.

    ``start`` is where the data begins in the padded axis (``overlap`` for the first
    tile, else 0), ``d`` the data extent and ``first`` whether the tile origin is 0.
    """
    m = np.full(T, -1, dtype=np.int64)
    j = np.arange(T)
    inside = (j >= start) & (j < start + d)
    m[inside] = j[inside] - start
    if first:
        src = 2 * start - j
        sel = (j < start) & (src < start + d)
        m[sel] = src[sel] - start
    elif d < T:
        src = 2 * (d - 1) - j
        sel = (j > d - 1) & (src >= 0)
        m[sel] = src[sel]
    return m


def z_map(sz, zpad):
    """Closed form of the two z reflections (``deconvolver.py:146-148``): padded plane -> data plane or -1."""
    Zp = sz + 2 * zpad
    m = np.full(Zp, -1, dtype=np.int64)
    m[zpad:zpad + sz] = np.arange(sz)
    for j in range(zpad):                       # 'in' about zpad, upper pad still zero
        s = 2 * zpad - j
        if s < Zp:
            m[j] = m[s] if s < zpad + sz else -1
    top = zpad + sz - 1                         # 'out' about zpad+sz-1, reads the filled lower pad
    for j in range(top + 1, Zp):
        s = 2 * top - j
        if s >= 0:
            m[j] = m[s]
    return m


def tile_wise(image, W, flat_field=None, blur_radius=None, tile_size=300, overlap=89, zpad=0, mode="fp64",
              only=None):
    """``deconvolver.py:98-170``: yields ``(x, y, res[:, :xend, :yend], pad[zpad:-zpad, :xend, :yend])``.

    ``only`` (a set of tile origins) skips the other tiles; PSF pairing is unchanged.
    """
    if zpad < 1 or overlap < 1:
        raise ValueError("zpad and overlap must be >= 1")
    ft = np.float64 if mode == "fp64" else np.float32
    sz, sx, sy = image.shape
    T = tile_size + 2 * overlap
    Zp = sz + 2 * zpad
    ws = cycle(W) if len(W) == 1 else iter(W)
    flat = None
    if flat_field is not None:                  # deconvolver.py:112: 2D flats broadcast over z
        flat = np.asarray(flat_field)
        if flat.ndim != image.ndim:
            flat = flat[np.newaxis]
    for (x, y, x0, x1, y0, y1), w in zip(tiled(image.shape, tile_size, overlap), ws):
        if only is not None and (x, y) not in only:
            continue
        tile = image[:, x0:x1, y0:y1]
        zdim, xdim, ydim = tile.shape
        xs = overlap if x == 0 else 0
        ys = overlap if y == 0 else 0
        xend, yend = xdim, ydim
        pad = np.zeros((Zp, T, T), dtype=ft)
        pad[zpad:zpad + zdim, xs:xs + xdim, ys:ys + ydim] = tile
        mask = (slice(zpad, Zp - zpad), slice(xs, xs + xdim), slice(ys, ys + ydim))
        if flat is not None:
            f = flat[:, x0:x1, y0:y1].astype(ft)
            pad[mask] = pad[mask] / f
        if x == 0:
            ofill.reflect(pad, overlap, mode="in", axis=1, out=pad)
            xend += overlap
        elif xdim < T:
            ofill.reflect(pad, xdim - 1, mode="out", axis=1, out=pad)
            xend += overlap
        if y == 0:
            ofill.reflect(pad, overlap, mode="in", axis=2, out=pad)
            yend += overlap
        elif ydim < T:
            ofill.reflect(pad, ydim - 1, mode="out", axis=2, out=pad)
            yend += overlap
        ofill.reflect(pad, zpad, mode="in", axis=0, out=pad)
        ofill.reflect(pad, zpad + zdim - 1, mode="out", axis=0, out=pad)
        spec = _fftn(pad)
        if mode != "fp64":
            spec = spec.astype(np.complex64)
        spec *= w
        res = _ifftn(spec)[zpad:Zp - zpad].real.astype(ft)
        if blur_radius is not None:
            b1 = oblur.box_1d(res, blur_radius, axis=1, dtype=ft)
            b2 = oblur.box_1d(b1, blur_radius, axis=2, dtype=ft)
            res = (res - b2).astype(ft)
        if flat is not None:
            pad[mask] = pad[mask] * f
        yield x, y, res[:, :xend, :yend], pad[zpad:Zp - zpad, :xend, :yend]


def apply(image, W, flat_field=None, blur_radius=None, tile_size=300, overlap=89, zpad=0, mode="fp64"):
    """``deconvolver.py:172-184``: stitch the owned ``[ov:-ov]`` part of every tile."""
    ft = np.float64 if mode == "fp64" else np.float32
    out = np.empty(image.shape, dtype=ft)
    for x, y, t, _ in tile_wise(image, W, flat_field, blur_radius, tile_size, overlap, zpad, mode):
        _, xd, yd = t.shape
        xd -= 2 * overlap
        yd -= 2 * overlap
        out[:, x:x + xd, y:y + yd] = t[:, overlap:-overlap, overlap:-overlap]
    return out


def full_deconv(image, psfs, flat_field=None, tile_size=300, zpad=None, overlap=1, beta=0.001, mode="fp64"):
    """``deconvolver.py:272-287``."""
    sz = image.shape[0]
    if zpad is None:
        zpad = sz
    W = wiener(psfs, sz, tile_size, zpad, overlap, beta, mode)
    return apply(image, W, flat_field, None, tile_size, overlap, zpad, mode)


def hyb_spots(image, W, flat_field, tile_size, overlap, zpad, blur_radius, threshold, delta, delta_fit,
              sigmaZ, sigmaXY, mode="fp64", tile_filter=None):
    """Hyb tile loop of ``__main__.py:254-265``: per-tile maxima+fit, keep-filter on the FITTED
    x/y, offset to FOV coordinates, stacked.  Returns ``(N, 8)``.

    ``tile_filter`` (a set of tile origins) may restrict the loop to a subset of tiles (used for
    bounded baselines).
    """
    ft = np.float64 if mode == "fp64" else np.float32
    out = []
    for x, y, tile, raw in tile_wise(image, W, flat_field, blur_radius, tile_size, overlap, zpad, mode,
                                     only=tile_filter):
        Xh = omax.find_local_maxima(tile, threshold=threshold, delta=delta, delta_fit=delta_fit, raw=raw,
                                    sigmaZ=sigmaZ, sigmaXY=sigmaXY, dtype=ft)
        lim = np.array([tile.shape[1] - overlap, tile.shape[2] - overlap])
        keep = np.all((Xh[:, 1:3] >= overlap) & (Xh[:, 1:3] < lim), axis=-1)
        Xh = Xh[keep]
        Xh[:, 1] += ft(x - overlap)
        Xh[:, 2] += ft(y - overlap)
        out.append(Xh)
    return np.vstack(out) if out else np.zeros((0, 8), dtype=ft)


def dapi_spots(image, W, flat_field, tile_size, overlap, zpad, blur_radius, threshold, delta, delta_fit,
               sigmaZ, sigmaXY, mode="fp64"):
    """DAPI path of ``__main__.py:217-225``: apply, RMS-normalise, maxima of +buffer and of -buffer."""
    buf = apply(image, W, flat_field, blur_radius, tile_size, overlap, zpad, mode)
    rms = np.sqrt(np.sum(buf.astype(np.float64) ** 2)) / np.sqrt(buf.size)
    ft = buf.dtype.type
    buf = (buf / ft(rms)).astype(buf.dtype)
    kw = dict(threshold=threshold, delta=delta, delta_fit=delta_fit, sigmaZ=sigmaZ, sigmaXY=sigmaXY, dtype=ft)
    plus = omax.find_local_maxima(buf, raw=image, **kw)
    minus = omax.find_local_maxima(-buf, raw=image, **kw)
    return plus, minus
