"""Host feed and output formats around the device path (reference ``mermake/io.py``).

* ``read_zarr`` decodes a zarr v2 directory array (the microscope's ``<fov>/data``): the
  reference reads it through dask (``io.py:124-126``); here the chunks are decoded with the
  standard library codecs (bz2, zlib, gzip, lzma) on a thread pool.
* ``read_im`` follows ``io.py:116-146``: frame 0 dropped, channels de-interleaved by the
  ``z_offsets`` tag of the XML, uint8 data squared into uint16.
* ``load_flats`` follows ``io.py:42-54`` on the device: centre crop, ``box((20, 20))`` through
  the HIP box blur, division by the median.
* ``FovFeeder`` reads FOVs ahead on a host thread and moves them to the device through two
  pinned staging buffers on a copy stream of its own (the prefetch of ``ImageQueue``,
  ``io.py:321-463``, with the host-to-device copy overlapped as well).
* ``OutputNames``, ``save_xfits`` and ``namespace_to_array`` write the reference's result
  files (``io.py:535-586``, ``io.py:619-637``): ``.npz`` with ``Xh`` or ``Xh_plus`` /
  ``Xh_minus``, ``version`` and ``args``.
"""
import bz2
import concurrent.futures
import glob
import gzip
import itertools
import json
import lzma
import os
import queue
import re
import threading
import zlib
from argparse import Namespace
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import torch

from . import __version__
from . import _lib as L
from . import blur as _blur
from ._dev import as_device

_CODECS = {
    "bz2": bz2.decompress,
    "zlib": zlib.decompress,
    "gzip": gzip.decompress,
    "lzma": lzma.decompress,
}


def _decoder(meta):
    comp = meta.get("compressor")
    if comp is None:
        return lambda b: b
    cid = comp.get("id")
    if cid not in _CODECS:
        raise NotImplementedError(f"zarr compressor {cid!r} is not supported (standard-library codecs: "
                                  f"{sorted(_CODECS)})")
    return _CODECS[cid]


def _workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def read_zarr(path, first=0, workers=None):
    """The zarr v2 array at directory ``path`` as a NumPy array, from index ``first`` of axis 0
    on (only the chunks that hold those frames are decoded).  Missing chunks read as
    ``fill_value``; filters are not supported."""
    with open(os.path.join(path, ".zarray")) as f:
        meta = json.load(f)
    if meta.get("zarr_format") != 2:
        raise ValueError(f"{path}: not a zarr v2 array")
    if meta.get("filters"):
        raise NotImplementedError(f"{path}: zarr filters are not supported")
    shape = tuple(meta["shape"])
    chunks = tuple(meta["chunks"])
    dtype = np.dtype(meta["dtype"])
    order = meta.get("order", "C")
    sep = meta.get("dimension_separator", ".")
    fill = meta.get("fill_value")
    dec = _decoder(meta)
    first = max(0, int(first))
    out_shape = (max(0, shape[0] - first),) + shape[1:] if shape else shape
    out = np.empty(out_shape, dtype=dtype)
    if fill is not None:
        out.fill(fill)
    elif out.size:
        out.fill(0)
    if out.size == 0:
        return out
    grid = [range((s + c - 1) // c) for s, c in zip(shape, chunks)]
    grid[0] = range(first // chunks[0], len(grid[0]))

    def one(idx):
        fn = os.path.join(path, sep.join(str(i) for i in idx))
        if not os.path.exists(fn):
            return
        with open(fn, "rb") as f:
            raw = dec(f.read())
        a = np.frombuffer(raw, dtype=dtype)
        if a.size != int(np.prod(chunks)):
            raise ValueError(f"{fn}: {a.size} elements, chunk holds {int(np.prod(chunks))}")
        a = a.reshape(chunks, order=order)
        lo = [i * c for i, c in zip(idx, chunks)]
        hi = [min(l + c, s) for l, c, s in zip(lo, chunks, shape)]
        src = [slice(0, h - l) for l, h in zip(lo, hi)]
        dst = [slice(l, h) for l, h in zip(lo, hi)]
        s0 = max(lo[0], first)  # frames before `first` are not returned
        src[0] = slice(s0 - lo[0], hi[0] - lo[0])
        dst[0] = slice(s0 - first, hi[0] - first)
        out[tuple(dst)] = a[tuple(src)]

    idxs = list(itertools.product(*grid))
    with concurrent.futures.ThreadPoolExecutor(workers or _workers()) as ex:
        for _ in ex.map(one, idxs):
            pass
    return out


def _xml_tag(txt, tag):
    return txt.split(tag)[-1].split("</")[0]


def read_im(path, return_pos=False, square=True):
    """``io.py:116-146``.  ``path`` names ``<dir>/<prefix>_<fov>.zarr``; the array is
    ``<dir>/<fov>/data``.  Returns (channels, z, x, y) when the XML beside it gives the channel
    count, else (frames, x, y); with ``return_pos`` also the stage position (x, y).
    ``square=False`` leaves uint8 frames unsquared (for ``SquaredU8`` stacks, squared on the
    device)."""
    dirname = os.path.dirname(path)
    fov = os.path.basename(path).split("_")[-1].split(".")[0]
    image = read_zarr(os.path.join(dirname, fov, "data"), first=1)
    shape = image.shape
    xml_file = os.path.splitext(path)[0] + ".xml"
    x = y = None
    if os.path.exists(xml_file):
        with open(xml_file, "r", encoding="ISO-8859-1") as f:
            txt = f.read()
        zstack = _xml_tag(txt, '<z_offsets type="string">')
        # the reference evaluates "x,y"; the tag holds two numbers
        x, y = (float(v) for v in _xml_tag(txt, '<stage_position type="custom">').split(","))
        nch = int(zstack.split(":")[-1])
        nzs = (shape[0] // nch) * nch
        image = image[:nzs].reshape([shape[0] // nch, nch, shape[-2], shape[-1]]).swapaxes(0, 1)
    if image.dtype == np.uint8 and square:  # squared in place of the widened copy
        image = image.astype(np.uint16)
        np.multiply(image, image, out=image)
    if return_pos:
        if x is None:
            raise ValueError(f"{path}: no XML with the stage position")
        return image, x, y
    return image


def center_crop(A, shape):
    """``io.py:34-40``: the centre (h, w) of the last two axes."""
    h, w = shape[-2:]
    H, W = A.shape[-2:]
    top = (H - h) // 2
    left = (W - w) // 2
    return A[top:top + h, left:left + w]


def _median(v):
    """``cupy.median`` of a float32 device tensor (``mm_median``): the middle value, or for an even
    count the float32 mean of the two middle values; a 0-d device tensor."""
    v = v.reshape(-1).contiguous()
    out = torch.empty((), dtype=torch.float32, device=v.device)
    ws = torch.empty(L.lib().mm_median_ws_bytes(), dtype=torch.uint8, device=v.device)
    L.call("mm_median", L.ptr(v), v.numel(), L.ptr(out), L.ptr(ws), ws.numel(), L.stream_ptr(v.device))
    return out


def load_flats(flat_field_tag, shape=None, device=None, **kwargs):
    """``io.py:42-54``: one flat per file ``flat_field_tag*`` (sorted), key ``im``; centre-crop
    to ``shape``, ``box((20, 20))`` over the last two axes, divide by the median.  Returns a
    (n, h, w) float32 device tensor."""
    stack = []
    for file in sorted(glob.glob(flat_field_tag + "*")):
        with np.load(file) as f:
            im = f["im"]
        if shape is not None:
            assert im.shape[-1] >= shape[-1] and im.shape[-2] >= shape[-2], "flat field smaller than image"
            im = center_crop(im, shape)
        cim = as_device(np.ascontiguousarray(im, dtype=np.float32), device)
        blurred = _blur.box(cim, (20, 20), axes=(-1, -2))
        blurred = blurred / _median(blurred)
        stack.append(blurred)
    if not stack:
        raise FileNotFoundError(f"no flat field files match {flat_field_tag}*")
    return torch.stack(stack)


class FovFeeder:
    """Iterate ``(path, channels, x, y)`` over FOV files with the channels already on the
    device: a host thread reads (decodes) up to ``prefetch`` FOVs ahead and copies each into
    one of two pinned staging buffers, and the buffer goes to the device on the feeder's copy
    stream.  The consumer's current stream waits for that copy (an event) before the channels
    are handed out, so they can go straight into ``HybPipeline`` / ``HybServer``.

    The reference's ``ImageQueue`` prefetches decoded images on a thread (``io.py:448-463``);
    its host-to-device copy then runs on the compute stream.  ``u8=True`` hands uint8 frames out
    as ``SquaredU8`` channels (squared by the tile loaders): half the bytes to stage and copy."""

    def __init__(self, paths, device=None, prefetch=2, u8=False):
        self.paths = list(paths)
        self.u8 = u8
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self._q = queue.Queue(maxsize=max(1, prefetch))
        self._stop = threading.Event()
        self._staging = [None, None]
        self._busy = [None, None]  # copy-done events of the staging buffers
        self._next = 0
        self.read_seconds = 0.0
        self.h2d_bytes = 0
        self._t = threading.Thread(target=self._work, daemon=True)
        self._t.start()

    def _work(self):
        import time
        for p in self.paths:
            if self._stop.is_set():
                break
            try:
                t0 = time.perf_counter()
                im, x, y = read_im(p, return_pos=True, square=not self.u8)
                self.read_seconds += time.perf_counter() - t0
                self._q.put((p, im, x, y, None))
            except Exception as e:  # reported to the consumer, as the reference warns and goes on
                self._q.put((p, None, None, None, e))
        self._q.put(None)

    def _stage(self, im):
        k = self._next
        self._next ^= 1
        if self._busy[k] is not None:
            self._busy[k].synchronize()  # the buffer's previous copy has left it
        nbytes = im.nbytes
        buf = self._staging[k]
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
            self._staging[k] = buf
        # one host pass: the (possibly strided, de-interleaved) image into the pinned buffer
        buf.numpy()[:nbytes].view(im.dtype).reshape(im.shape)[...] = im
        with torch.cuda.stream(self.copy_stream):
            raw = buf[:nbytes].to(self.device, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
        if im.dtype == np.uint8:
            from .deconvolver import SquaredU8
            dev = SquaredU8.of(raw.view(im.shape))
        else:
            dev = raw.view(torch.uint16 if im.dtype == np.uint16 else torch.float32).view(im.shape)
        self._busy[k] = ev
        self.h2d_bytes += nbytes
        return dev, ev

    def __iter__(self):
        return self

    def __next__(self):
        item = self._q.get()
        if item is None:
            raise StopIteration
        p, im, x, y, err = item
        if err is not None:
            raise RuntimeError(f"failed to read {p}: {err}") from err
        if im.dtype not in (np.uint16, np.float32) and not (self.u8 and im.dtype == np.uint8):
            im = im.astype(np.float32)
        dev, ev = self._stage(im)
        torch.cuda.current_stream(self.device).wait_event(ev)
        dev.record_stream(torch.cuda.current_stream(self.device))
        chans = [dev[i] for i in range(dev.shape[0])] if dev.ndim == 4 else [dev]
        return p, chans, x, y

    def close(self):
        self._stop.set()
        try:
            while True:
                self._q.get_nowait()
        except queue.Empty:
            pass
        self._t.join(timeout=2.0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ----------------------------------------------------------------- output files
def get_ih(file_path):
    """``io.py:149-155``: the first number in the base name (10**100 if none)."""
    m = re.search(r"\d+", os.path.basename(file_path))
    return int(m.group()) if m else 10 ** 100


def get_ifov(file_path):
    """``io.py:157-162``: the last digits before ``.zarr``."""
    m = re.search(r"([0-9]+)[^0-9]*\.zarr", Path(file_path).name)
    if m:
        return int(m.group(1))
    raise ValueError(f"No digits found before .zarr in filename: {Path(file_path).name}")


def get_iset(file_path):
    """``io.py:164-180``: the digits after ``_set`` in the file name or the nearest parent."""
    path = Path(file_path)
    m = re.search(r"_set([0-9]+)", path.name)
    if m:
        return int(m.group(1))
    if path.parent == path:
        raise ValueError(f"No digits found after the word _set in {file_path}")
    return get_iset(path.parent)


class OutputNames:
    """File names of ``ImageQueue.get_name`` (``io.py:535-546``); the templates default to
    those of the reference's ``example.toml``."""

    def __init__(self, hyb_save="{fov}--{tag}--col{icol}__Xhfits.npz", dapi_save="{fov}--{tag}--dapiFeatures.npz",
                 drift_save="drift_{fov}--_set{iset}.pkl"):
        self.hyb_save = hyb_save
        self.dapi_save = dapi_save
        self.drift_save = drift_save

    def get_name(self, path, icol=None):
        p = Path(path)
        fov, tag = p.stem, p.parent.name
        if icol is None:
            return self.drift_save.format(ifov=get_ifov(path), fov=fov, iset=get_iset(path))
        if icol == -1:
            return self.dapi_save.format(fov=fov, tag=tag)
        return self.hyb_save.format(fov=fov, tag=tag, icol=icol)


def namespace_to_array(obj, prefix=""):
    """``io.py:619-637``: (block, key, value) string rows of a nested settings namespace."""
    rows = []
    if isinstance(obj, (Namespace, SimpleNamespace)):
        obj = vars(obj)
    if isinstance(obj, dict):
        for k, v in obj.items():
            full = f"{prefix}.{k}" if prefix else k
            if isinstance(v, (Namespace, SimpleNamespace, dict)):
                rows.extend(namespace_to_array(v, prefix=full))
            else:
                rows.append((prefix, k, str(v)))
    else:
        rows.append((prefix, "", str(obj)))
    return rows


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def save_xfits(filepath, data, args_array, version=__version__):
    """``ImageQueue.save_xfits`` (``io.py:562-586``): ``np.savez_compressed`` of ``data``
    (``{'Xh': ...}`` or ``{'Xh_plus': ..., 'Xh_minus': ...}``) with ``version`` and ``args``,
    unless the file exists.  Device arrays are copied to the host first.  Returns True when
    the file was written."""
    if os.path.exists(filepath):
        return False
    arrays = {k: _host(v) for k, v in data.items()}
    np.savez_compressed(filepath, version=version, args=np.asarray(args_array, dtype=str), **arrays)
    return True
