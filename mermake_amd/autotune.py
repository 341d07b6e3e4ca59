"""Autotuner of the tiled deconvolution's schedule: the FFT form of the T-length y and x
passes, the form of the z pass, and the tiles per launch.

The reference hands its transforms to cuFFT plans (``deconvolver.py:130-158``) and runs
one tile per call (``deconvolver.py:116-170``).  Here the transforms are this package's
own kernels, in several forms per pass (``Deconvolver.set_schedule``):

* y/x passes: ``stockham`` (LDS Stockham radix passes), ``prime_first`` (a
  matrix-product stage of the prime p of T = p * M, then radix M), ``rader`` (the
  prime 113 or 239 by Rader's algorithm, Good-Thomas, column-blocked spectra);
* z pass: ``fft`` (z FFTs with the Wiener multiply), ``matrix`` (the folded
  z-kernel matrix), ``dft80`` (length-80 DFTs, zpad = sz - 1);
* tiles per launch: the memory bound, or a half or a quarter of it (the most tiles within 2 % of the
  least time).

``tune`` times, on the device, every admitted (y/x, z) pair on a batch of tiles of
the caller's channel shape (a seeded synthetic uint16 stack), keeps the fastest, then
times the tiles per launch with it.  Every form computes the same deconvolution
(relative L2 within 1e-5 of the float64 result); they differ in speed.  Results are
cached per device, library build and geometry in a JSON file
(``MERMAKE_AMD_TUNE_CACHE``, default ``~/.cache/mermake_amd/schedule.json``), so a
second process with the same geometry reads the schedule instead of timing it.
"""
import json
import os
import time

import torch

from . import _lib as L
from .deconvolver import tile_geometry, tile_table


def cache_path():
    return os.environ.get("MERMAKE_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mermake_amd", "schedule.json"))


def _key(d, channel_shape, ntiles, window):
    name = torch.cuda.get_device_name(d.device)
    try:
        stamp = int(os.path.getmtime(L.LIBPATH))
    except OSError:
        stamp = 0
    return "|".join([name, f"lib{L.lib().mm_version()}.{stamp}", f"sz{d.tile_height}", f"T{d._plan.T}", f"Zp{d._plan.Zp}",
                     f"npsf{'1' if d.psf_fft.shape[0] == 1 else 'n'}", f"tiles{ntiles}",
                     f"win{window[0]}-{window[1]}" if window else "win-"])


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, key, entry):
    try:
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        data = _load(path)
        data[key] = entry
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass   # a read-only home keeps the schedule for this process only


def _stack(shape, device, seed=0):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    v = torch.randint(0, 4096, tuple(shape), generator=g, dtype=torch.int32, device=device)
    return v.to(torch.uint16)


def _run(d, table, ntiles, B, window):
    """Deconvolve ``ntiles`` descriptors in launches of ``B`` tiles (as ``HybPipeline`` batches them)."""
    single = d.psf_fft.shape[0] == 1
    rec = L.TILE_DTYPE.itemsize
    for b0 in range(0, ntiles, B):
        nb = min(B, ntiles - b0)
        d.run_tiles(table[b0 * rec:(b0 + nb) * rec], nb, nb if single else 1, None, window=window)


def _time(d, table, ntiles, B, window, reps):
    """Milliseconds of the whole batch, the least of ``reps`` runs after a warm-up run."""
    st = torch.cuda.current_stream(d.device)
    _run(d, table, ntiles, B, window)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        _run(d, table, ntiles, B, window)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return best


def tune(d, channel_shape, stacks=1, window=None, reps=3, cache=True, seed=0):
    """Time the schedules of ``d`` (a ``Deconvolver``) on the tiles of ``stacks`` channel stacks of
    ``channel_shape`` and select the fastest (``d.set_schedule``).  ``window=(lo, hi)`` times the
    windowed deconvolution the hyb caller runs (``mm_deconv_tiles_window``).  Returns the record kept
    in ``d.tuning``: the chosen forms and tiles per launch, and the milliseconds of every candidate."""
    sz, sx, sy = (int(v) for v in channel_shape)
    if sz != d.tile_height:
        raise ValueError(f"channel depth {sz} does not match the Deconvolver's {d.tile_height}")
    geom = tile_geometry((sz, sx, sy), d.tile_size, d.overlap, d.psf_fft.shape[0], d.psf_keys, d.psf_mode)
    ntiles = len(geom) * max(1, int(stacks))
    path = cache_path()
    key = _key(d, (sz, sx, sy), ntiles, window)
    if cache:
        hit = _load(path).get(key)
        if hit is not None:
            sched = d.schedule()
            if hit.get("xy") in sched["xy_forms"] and hit.get("z") in sched["z_forms"]:
                d.set_schedule(hit["xy"], hit["z"])
                d.batch_tiles = hit.get("batch_tiles")
                d.tuning = dict(hit, cached=True)
                return d.tuning
    t0 = time.perf_counter()
    img = _stack((sz, sx, sy), d.device, seed)
    tab1 = tile_table(geom, img, None)
    table = tab1 if stacks <= 1 else torch.cat([tab1] * int(stacks))
    sched = d.schedule()
    d.batch_tiles = None
    B0 = d.batch_size(ntiles)
    forms = {}
    for xy in sched["xy_forms"]:
        for z in sched["z_forms"]:
            d.set_schedule(xy, z)
            forms[f"{xy}+{z}"] = _time(d, table, ntiles, B0, window, reps)
    best = min(forms, key=forms.get)
    xy, z = best.split("+")
    d.set_schedule(xy, z)
    batches = {}
    for B in sorted({B0, max(1, -(-B0 // 2)), max(1, -(-B0 // 4))}, reverse=True):
        batches[str(B)] = forms[best] if B == B0 else _time(d, table, ntiles, B, window, reps)
    # the most tiles per launch within 2 % of the least time: fewer launches and host synchronisations
    # for the callers, and no choice made by timing noise
    fastest = min(batches.values())
    B = max(int(b) for b, ms in batches.items() if ms <= 1.02 * fastest)
    d.batch_tiles = None if B == B0 else B   # None: as many tiles per launch as memory allows
    del img, table, tab1
    d.tuning = {"xy": xy, "z": z, "batch_tiles": d.batch_tiles, "tiles": ntiles, "ms_forms": forms,
                "ms_batch_tiles": batches,
                "tune_s": round(time.perf_counter() - t0, 3), "cached": False}
    if cache:
        _store(path, key, {k: v for k, v in d.tuning.items() if k != "cached"})
    return d.tuning
