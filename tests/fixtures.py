"""Loaders of the committed golden vectors (tests/golden/) and writers of synthetic zarr FOVs."""
import bz2
import json
import os
import zlib

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TILES = ((0, 0), (300, 300), (900, 900))
TS, OV = 300, 89


def fixture_windows():
    """{(x, y): (bounds, uint16 window (read_im squares the uint8 frames), golden f16 planes)}, golden_z, shape."""
    f = np.load(os.path.join(GOLDEN, "fixture_tiles.npz"))
    out = {}
    for (x, y) in TILES:
        shp = tuple(f[f"winshape_{x}_{y}"])
        w8 = np.frombuffer(bz2.decompress(f[f"winbz2_{x}_{y}"].tobytes()), np.uint8).reshape(shp)
        out[(x, y)] = (tuple(int(v) for v in f[f"bounds_{x}_{y}"]), w8.astype(np.uint16) ** 2, f[f"gold_{x}_{y}"])
    return out, tuple(int(z) for z in f["golden_z"]), tuple(int(s) for s in f["shape"])


def fixture_stack():
    """35x1200x1200 uint16 stack holding the three windows (zeros elsewhere)."""
    wins, gz, shape = fixture_windows()
    stack = np.zeros(shape, dtype=np.uint16)
    for (x0, x1, y0, y1), w, _ in wins.values():
        stack[:, x0:x1, y0:y1] = w
    return stack, wins, gz


def psf_750():
    return np.load(os.path.join(GOLDEN, "psf_750.npy"))


def dapi_features():
    f = np.load(os.path.join(GOLDEN, "dapi_features.npz"))
    return f["Xh_plus"], f["Xh_minus"]


def drift_known_answer():
    return json.load(open(os.path.join(GOLDEN, "drift_known_answer.json")))


def golden_accept(out_f32, gold_f16):
    """Per-voxel acceptance against the reference's float16 golden output.

    The reference output is a float32 computation rounded to float16, so the
    oracle/kernel result must round to the same float16 on nearly every voxel;
    elsewhere it may differ by float32 round-off of the reference (absolute,
    relative to the volume's peak) plus one float16 ulp.
    Returns (fraction exactly equal, worst violation ratio).
    """
    g = gold_f16.astype(np.float64)
    o16 = out_f32.astype(np.float16).astype(np.float64)
    exact = float((o16 == g).mean())
    ulp = np.spacing(np.abs(gold_f16)).astype(np.float64)
    allow = ulp + 2e-6 * float(np.abs(g).max())
    worst = float((np.abs(out_f32.astype(np.float64) - g) / allow).max())
    return exact, worst


def write_zarr(path, a, chunks, compressor="bz2", order="C", sep=".", fill_value=0, skip=()):
    """A zarr v2 directory array, written from the format's definition: one file per chunk,
    key = chunk indices joined by ``sep``, edge chunks stored full size (padded with
    fill_value), chunks in ``skip`` left out."""
    os.makedirs(path, exist_ok=True)
    meta = {"zarr_format": 2, "shape": list(a.shape), "chunks": list(chunks), "dtype": a.dtype.str,
            "compressor": None if compressor is None else {"id": compressor}, "fill_value": fill_value,
            "filters": None, "order": order}
    if sep != ".":
        meta["dimension_separator"] = sep
    json.dump(meta, open(os.path.join(path, ".zarray"), "w"))
    grid = [range((s + c - 1) // c) for s, c in zip(a.shape, chunks)]
    for idx in np.ndindex(*[len(g) for g in grid]):
        if idx in skip:
            continue
        blk = np.full(chunks, fill_value, dtype=a.dtype)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, a.shape))
        blk[tuple(slice(0, x.stop - x.start) for x in sl)] = a[sl]
        raw = blk.tobytes(order=order)
        if compressor == "bz2":
            raw = bz2.compress(raw, 1)
        elif compressor == "zlib":
            raw = zlib.compress(raw, 1)
        fn = os.path.join(path, sep.join(str(i) for i in idx))
        os.makedirs(os.path.dirname(fn), exist_ok=True)
        open(fn, "wb").write(raw)


def make_fov(root, fov, frames, nch=3, pos="-4351.10,5862.40"):
    """A microscope FOV as ``read_im`` expects it: ``<root>/<prefix>_<fov>.zarr`` plus
    ``<root>/<fov>/data`` (one frame per chunk) and the XML beside the ``.zarr``."""
    os.makedirs(root, exist_ok=True)
    path = os.path.join(root, f"Conv_zscan__{fov}.zarr")
    os.makedirs(path, exist_ok=True)
    write_zarr(os.path.join(root, fov, "data"), frames, (1,) + frames.shape[1:], "bz2")
    xml = ('<?xml version="1.0" encoding="ISO-8859-1"?>\n<settings>\n'
           f'  <stage_position type="custom">{pos}</stage_position>\n'
           f'  <z_offsets type="string">0:-15:-0.4:{nch}</z_offsets>\n</settings>\n')
    open(os.path.join(root, f"Conv_zscan__{fov}.xml"), "w").write(xml)
    return path
