"""Benchmark of the MERMAKE image hot path on one MI355X.

Workload (one step = one FOV): the full hyb FOV, 3 colour channels of
40 x 2800 x 2800 uint16 with a float32 flat field, one 40x64x64 Gaussian PSF,
tile_size 500, overlap 89 (36 padded 678^2 tiles per stack), beta 1e-4,
blur_radius 30, threshold 3600, delta 1, delta_fit 3, zpad = sz - 1.  Each
channel goes through deconvolution -> blur subtraction -> local maxima ->
delta_fit -> keep-filter, batched over its 36 tiles.  Synthetic seeded data
(puncta density 1/20^3, Poisson background 500, quadratic illumination).
The timed FOV steps run as a stream (``HybPipeline.submit``, two FOVs in
flight: each FOV's chain is enqueued while the previous one runs); the same
steps one at a time are reported beside them (``sequential``).

Prints one JSON line: FOV voxels/s, the roofline of the dominant kernel
(HIP-event timed on the stream the kernel runs on), the CPU baseline (the
float32 NumPy/SciPy restatement on a bounded tile sample, extrapolated), and
the plain one-tile-at-a-time GPU form for comparison.  Further legs (``legs``)
measure the other workloads: the 100-PSF grid (40x3000^2 at tile 300) with its
DAPI pass on nuclei, a batch of uneven FOVs (40 x {2048, 2304, 2800}^2) with
DAPI-pair drift, and the served stream (2 hyb + 1 DAPI channels per request,
single or multi-PSF, block drift) at Poisson rates against the plain form.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK = 8.0e12
CFG = dict(sz=40, sx=2800, sy=2800, channels=3, tile_size=500, overlap=89, beta=1e-4, blur_radius=30,
           threshold=3600, delta=1, delta_fit=3, sigmaZ=1, sigmaXY=1.5, psf_shape=(40, 64, 64), psf_sigma=(1.5, 1.2),
           density=1.0 / 20 ** 3)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class mark:
    """A named range on the profiler's marker trace (roctx through torch.cuda.nvtx; free when no profiler
    listens), so that kernel traces can be split by benchmark phase."""

    def __init__(self, name):
        self.name = name

    def __enter__(self):
        import torch
        torch.cuda.nvtx.range_push(self.name)

    def __exit__(self, *exc):
        import torch
        torch.cuda.nvtx.range_pop()


def m1_bytes_per_stack(geom, sz, T, Zp, flat=True):
    """Algorithmic-bytes model M1 per channel stack (one HBM round trip per stage)."""
    H = T // 2 + 1
    tot = 0
    for g in geom:
        tot += 2 * sz * g["xd"] * g["yd"] + (4 * g["xd"] * g["yd"] if flat else 0) + 8 * Zp * T * H \
            + 2 * 8 * Zp * T * H + 2 * 4 * sz * g["xe"] * g["ye"]
    return tot


def pass_bytes(geom, sz, T, wplanes, window=(None, None), single_psf=True, flat=True):
    """Compulsory HBM bytes of each pass of mm_deconv_tiles for one batch (the kernels' algorithmic traffic).

    wplanes: planes of the z-pass operator (240 for the length-80 z pass); window: the rows and columns
    [lo, hi) of the deconvolved tiles that the C2R produces (None: whole tiles)."""
    H = T // 2 + 1
    n = len(geom)
    spec = 8 * sz * T * H
    inp = sum(2 * sz * g["xd"] * g["yd"] + (4 * g["xd"] * g["yd"] if flat else 0) for g in geom)
    W = 8 * wplanes * T * H * (1 if single_psf else n)
    lo, hi = window
    rows = T if lo is None else hi - lo
    return {"r2c": inp + n * spec, "x_fwd": 2 * n * spec, "z_wiener": 2 * n * spec + W, "x_inv": 2 * n * spec,
            "c2r": n * spec * rows // T + n * 4 * sz * rows * rows}


def deconv_window(pipe, T):
    """Rows and columns [lo, hi) of the deconvolved tiles the hyb pipeline computes (the blur window, grown by the
    blur half width unless the blur rides on the operator, in whole C2R row pairs), or (None, None)."""
    win = pipe.deconv_window()
    if win is None:
        return None, None
    return win[0] & ~1, min(T, (win[1] + 1) & ~1)


PASS_KERNEL = {"r2c": "k_rd_r2c", "x_fwd": "k_rd_c2c<false", "z_wiener": "k_tile_zfft", "x_inv": "k_rd_c2c<true",
               "c2r": "k_rd_c2r"}


def measured_traffic(pass_name):
    """HBM bytes per launch of the pass's kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/hyb_fov_hbm_traffic.md), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hyb_fov_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("roofline-traffic:"):
                _, name, val = line.split()
                if name == PASS_KERNEL.get(pass_name):
                    return float(val)
    except OSError:
        return None
    return None


def cpu_baseline(stack_np, flat_np, psf, cfg, ntiles_per_stack):
    """Oracle (float32 mode) on a tile sample: corner, two edges, interior of one channel."""
    from oracle import deconv as odec
    sample = {(0, 0), (0, 1000), (2500, 1000), (1000, 1000)}
    zpad = cfg["sz"] - 1
    W = odec.wiener(psf, cfg["sz"], cfg["tile_size"], zpad, cfg["overlap"], cfg["beta"], mode="fp32")
    kw = dict(blur_radius=cfg["blur_radius"], threshold=cfg["threshold"], delta=cfg["delta"],
              delta_fit=cfg["delta_fit"], sigmaZ=cfg["sigmaZ"], sigmaXY=cfg["sigmaXY"], mode="fp32")
    odec.hyb_spots(stack_np, W, flat_np, cfg["tile_size"], cfg["overlap"], zpad, tile_filter={(1000, 1000)}, **kw)
    t0 = time.perf_counter()
    odec.hyb_spots(stack_np, W, flat_np, cfg["tile_size"], cfg["overlap"], zpad, tile_filter=sample, **kw)
    per_tile = (time.perf_counter() - t0) / len(sample)
    fov_vox = cfg["channels"] * cfg["sz"] * cfg["sx"] * cfg["sy"]
    fov_s = per_tile * ntiles_per_stack * cfg["channels"]
    cores = len(os.sched_getaffinity(0))
    return {"value": fov_vox / fov_s, "unit": "voxels/s", "cores": cores, "kind": "port",
            "sample": f"oracle float32 (NumPy/SciPy, scipy.fft workers={cores}) hyb chain on 4 tiles "
                      f"(corner, 2 edges, interior) of channel 0, {per_tile:.2f} s/tile, extrapolated to "
                      f"{ntiles_per_stack} tiles x {cfg['channels']} channels"}


def warm_copies(srv, tensors, dev, times=2):
    """Warm this process's host-to-device path on the server's copy stream with the stream's own pinned
    inputs: on a fresh GPU the first several GB copied move at about half the DMA rate."""
    import torch
    with torch.cuda.stream(srv.copy_stream):
        for t in tensors:
            for _ in range(times):
                torch.empty(t.shape, dtype=t.dtype, device=dev).copy_(t, non_blocking=True)
    torch.cuda.synchronize()


def serve_leg(pipe, stacks, flat, cfg, args, dev):
    """The request loop saturated with hyb FOVs of mixed sizes (40 x {2048, 2304, 2800}^2, 3 channels each, host
    inputs in pinned memory, all requests at t = 0): uint16 stacks, then the uint8 frames whose squares are the
    pixels."""
    import torch
    from mermake_amd.deconvolver import SquaredU8
    from mermake_amd.serve import FovRequest, HybServer
    from mermake_amd.synth import stack_torch
    fovs, fovs_u8 = {}, {}
    for side in (2048, 2304, 2800):
        if side == cfg["sx"]:
            chans, fl = list(stacks), flat
        else:
            chans, fl = [], None
            for c in range(cfg["channels"]):
                s_, f_ = stack_torch(2000 + 10 * side + c, (cfg["sz"], side, side), cfg["density"], dev, flat=True)
                chans.append(s_)
                fl = f_
        fovs[side] = ([c.cpu().pin_memory() for c in chans], fl.cpu().pin_memory())
        # the microscope's uint8 frames of the same shapes (read_im squares them): half the PCIe bytes
        fovs_u8[side] = ([SquaredU8.of(torch.sqrt(c.float()).clamp_(max=255).to(torch.uint8).cpu()).pin_memory()
                          for c in chans], fovs[side][1])
        torch.cuda.synchronize()
    sides = [2048, 2304, 2800]
    rng = np.random.default_rng(7)
    order = [sides[i] for i in rng.integers(0, 3, size=args.serve_requests)]
    srv = HybServer(pipe, batch_tiles=36, slots=2)
    out = {"fov_sides": order, "inputs": "host pinned uint16 stacks + float32 flat (H2D inside the loop); "
                                         "saturated_u8: the uint8 frames whose squares are the pixels (SquaredU8)"}
    for name, arr, src in (("saturated", np.zeros(len(order)), fovs),
                           ("saturated_u8", np.zeros(len(order)), fovs_u8)):
        srv.run([FovRequest(-1 - i, *src[side]) for i, side in enumerate(order[:2])])  # warm the slots
        warm_copies(srv, [t for side in sides for t in src[side][0]], dev)
        reqs = [FovRequest(i, *src[side], arrival=float(a)) for i, (side, a) in enumerate(zip(order, arr))]
        with mark(f"served_hyb_{name}"):
            srv.run(reqs)
        st = dict(srv.stats)
        vox = sum(cfg["channels"] * cfg["sz"] * s_ * s_ for s_ in order)
        st["voxels_per_s"] = vox / st["wall_s"]
        log(f"served {name}: {st['fovs_per_s']:.2f} FOV/s, p50 {st['p50_latency_s'] * 1e3:.1f} ms, "
            f"p99 {st['p99_latency_s'] * 1e3:.1f} ms")
        out[name] = st
    del srv
    return out


def feed_leg(dev, cfg):
    """Host feed rates: bz2 zarr decode (``read_im`` on a thread pool) of a sample of microscope
    frames, and the pinned host-to-device copy of one whole FOV of uint16 stacks."""
    import shutil
    import tempfile
    import torch
    from mermake_amd import io as mio
    side, nch, nz = cfg["sx"], cfg["channels"], 5  # 16 chunks: one per decode thread
    rng = np.random.default_rng(3)
    frames = rng.integers(18, 31, (1 + nch * nz, side, side), dtype=np.uint8)  # ~sqrt of 600-count Poisson frames
    root = tempfile.mkdtemp(prefix="mm_feed_")
    try:
        d = os.path.join(root, "H1_set1", "001", "data")
        os.makedirs(d)
        meta = {"zarr_format": 2, "shape": list(frames.shape), "chunks": [1, side, side], "dtype": "|u1",
                "compressor": {"id": "bz2", "level": 9}, "fill_value": 0, "filters": None, "order": "C"}
        json.dump(meta, open(os.path.join(d, ".zarray"), "w"))
        import bz2
        import concurrent.futures

        def put(i):
            open(os.path.join(d, f"{i}.0.0"), "wb").write(bz2.compress(frames[i].tobytes(), 9))
        with concurrent.futures.ThreadPoolExecutor(16) as ex:
            list(ex.map(put, range(frames.shape[0])))
        path = os.path.join(root, "H1_set1", "Conv_zscan__001.zarr")
        os.makedirs(path)
        open(os.path.splitext(path)[0] + ".xml", "w").write(
            f'<settings><stage_position type="custom">0,0</stage_position>'
            f'<z_offsets type="string">0:-15:-0.4:{nch}</z_offsets></settings>')
        mio.read_im(path)
        a = time.perf_counter()
        im = mio.read_im(path)
        dec_s = time.perf_counter() - a
        assert im.shape == (nch, nz, side, side)
        raw_mb = frames[1:].nbytes / 1e6
        fov_frames = 1 + nch * cfg["sz"]
        host = torch.empty((nch, cfg["sz"], side, side), dtype=torch.uint16, pin_memory=True)
        cs = torch.cuda.Stream(device=dev)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        with torch.cuda.stream(cs):
            host.to(dev, non_blocking=True)
            ev[0].record(cs)
            devt = host.to(dev, non_blocking=True)
            ev[1].record(cs)
        torch.cuda.synchronize()
        h2d_ms = ev[0].elapsed_time(ev[1])
        del host, devt
        workers = mio._workers()
        rate = raw_mb / dec_s
        out = {"sample": f"{frames.shape[0] - 1} bz2 (level 9) frames of {side}^2 uint8 ({nz} z x {nch} channels), "
                         f"read_im on {workers} threads",
               "decode_MB_per_s": rate, "threads": workers,
               "fov_decode_s_est": (fov_frames * side * side / 1e6) / rate,
               "h2d_GB_per_s": host_bytes(cfg) / (h2d_ms * 1e-3) / 1e9, "fov_h2d_ms": h2d_ms}
        log(f"feed: bz2 decode {rate:.0f} MB/s on {workers} threads (a {fov_frames}-frame FOV: "
            f"{out['fov_decode_s_est']:.2f} s), H2D {out['h2d_GB_per_s']:.1f} GB/s ({h2d_ms:.1f} ms per FOV)")
        return out
    finally:
        shutil.rmtree(root, ignore_errors=True)


def dapi_leg(stack, flat, psf, cfg):
    """The DAPI block of one FOV (``__main__.py:213-229``) on one 40x2800x2800 stack: deconvolution with
    zpad (sz - 1) // 2 and beta 0.01, blur 50, RMS normalisation, maxima and fits of both signs (delta 5,
    delta_fit 5)."""
    import torch
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.pipeline import DapiPipeline
    sz = cfg["sz"]
    d = Deconvolver(psf, tuple(stack.shape), tile_size=cfg["tile_size"], zpad=(sz - 1) // 2,
                    overlap=cfg["overlap"], beta=0.01)
    dp = DapiPipeline(d, blur_radius=50, threshold=3.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5)
    plus, minus = dp(stack, flat)
    torch.cuda.synchronize()
    reps = 2
    a = time.perf_counter()
    for _ in range(reps):
        plus, minus = dp(stack, flat)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - a) / reps * 1e3
    vox = stack.numel()
    log(f"DAPI stack: {ms:.1f} ms ({vox / (ms * 1e-3) / 1e9:.1f} G voxels/s), {plus.shape[0]} + {minus.shape[0]} spots")
    out = {"ms_per_stack": ms, "voxels_per_s": vox / (ms * 1e-3), "spots_plus": int(plus.shape[0]),
           "spots_minus": int(minus.shape[0]), "zpad": (sz - 1) // 2, "Zp": d._plan.Zp}
    del dp, d
    torch.cuda.empty_cache()
    return out


def grid_psfs(n, ts, shape=(40, 64, 64)):
    """n x n Gaussian PSFs keyed (0, x, y) at the centres of the tiles of size ts, x-major, widths varying
    smoothly across the field (sigma_xy 1.0 -> 1.6, sigma_z 1.2 -> 2.0)."""
    from mermake_amd.synth import gaussian_psf
    out = {}
    for i in range(n):
        for j in range(n):
            t = (i + j) / max(1, 2 * (n - 1))
            out[(0, i * ts + ts // 2, j * ts + ts // 2)] = gaussian_psf(shape, 1.2 + 0.8 * t, 1.0 + 0.6 * t)
    return out


def _timed(fn, reps):
    """Mean wall ms of ``fn`` over ``reps`` runs after one warm-up (synchronised on both sides)."""
    import torch
    out = fn()
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - a) / reps * 1e3, out


def multi_psf_leg(dev, cfg, grid_pipe, dapi_grid, reps=3):
    """The multi-PSF grid workload: one 40x3000^2 hyb stack at tile 300 / overlap 89 (100 tiles of T 478, Zp 118),
    each tile with its own of 100 PSFs; and the DAPI pass of the same shape and PSF grid (zpad 19, Zp 78,
    beta 0.01, blur 50, threshold 3, delta 5, delta_fit 5) on 2000 nuclei."""
    import torch
    from mermake_amd.deconvolver import tile_geometry
    from mermake_amd.synth import nuclei_torch, stack_torch
    sz, side = cfg["sz"], 3000
    stack, flat = stack_torch(7000, (sz, side, side), cfg["density"], dev, flat=True)
    d = grid_pipe.d
    geom = d.geometry(stack.shape)
    assert [g["psf"] for g in geom] == list(range(100)), "the 10x10 grid pairs each tile with its own PSF"
    ms, out = _timed(lambda: grid_pipe(stack, flat), reps)
    vox = stack.numel()
    m1 = m1_bytes_per_stack(geom, sz, d._plan.T, d._plan.Zp)
    hyb = {"workload": "40x3000^2 uint16 + flat, tile 300 / overlap 89, 100 tiles of 118x478x478 each with its own "
                       "of 100 PSFs (keyed 10x10 grid)", "ms_per_stack": ms, "voxels_per_s": vox / (ms * 1e-3),
           "m1_roofline_frac": m1 / (ms * 1e-3) / HBM_PEAK, "m1_bytes": m1, "spots": int(out.shape[0])}
    log(f"multi-PSF grid: {ms:.1f} ms per 40x3000^2 stack ({vox / ms / 1e6:.1f} G voxels/s, M1 "
        f"{hyb['m1_roofline_frac']:.3f}), {out.shape[0]} spots")
    del stack, flat, out
    nuc = nuclei_torch(7001, (sz, side, side), 2000, dev)
    ms, (pl, mi) = _timed(lambda: dapi_grid(nuc), reps)
    dd = dapi_grid.d
    m1 = m1_bytes_per_stack(tile_geometry(nuc.shape, dd.tile_size, dd.overlap), sz, dd._plan.T, dd._plan.Zp,
                            flat=False)
    dapi = {"workload": "DAPI 40x3000^2 uint16, 2000 nuclei, the same 100-PSF grid, tile 300 / overlap 89, zpad 19 "
                        "(Zp 78), beta 0.01, blur 50, threshold 3, delta 5, delta_fit 5", "ms_per_stack": ms,
            "voxels_per_s": vox / (ms * 1e-3), "m1_roofline_frac": m1 / (ms * 1e-3) / HBM_PEAK, "m1_bytes": m1,
            "spots_plus": int(pl.shape[0]), "spots_minus": int(mi.shape[0])}
    log(f"DAPI nuclei (100-PSF grid): {ms:.1f} ms per stack ({vox / ms / 1e6:.1f} G voxels/s, M1 "
        f"{dapi['m1_roofline_frac']:.3f}), {pl.shape[0]} + {mi.shape[0]} spots")
    for rec, dd_ in ((hyb, d), (dapi, dd)):
        t = dd_.tuning
        rec["schedule"] = ({k: t[k] for k in ("xy", "z", "batch_tiles", "ms_forms", "ms_batch_tiles", "tune_s")}
                           if t else dd_.schedule())
    del nuc
    torch.cuda.empty_cache()
    return {"hyb": hyb, "dapi": dapi}


def uneven_leg(dev, cfg, pipe, dapi_pipe, reps=3):
    """Batched uneven tiles + drift: three hyb stacks of 40 x 2048^2, 2304^2 and 2800^2 (interior and edge tiles of
    different sizes) in one run_many; and for each size a DAPI pair, the second shifted by a seeded integer offset
    (dz in +-3, dy, dx in +-30) with independent noise, whose drift must come back exactly."""
    import types
    import torch
    from mermake_amd.align import DualAligner
    from mermake_amd.deconvolver import tile_geometry
    from mermake_amd.synth import nuclei_torch, stack_torch
    sz = cfg["sz"]
    sides = (2048, 2304, 2800)
    stacks, flats = [], []
    for i, side in enumerate(sides):
        s, f = stack_torch(7100 + i, (sz, side, side), cfg["density"], dev, flat=True)
        stacks.append(s)
        flats.append(f)
    ms, out = _timed(lambda: pipe.run_many(stacks, flats), reps)
    T, Zp = pipe.d._plan.T, pipe.d._plan.Zp
    m1 = sum(m1_bytes_per_stack(tile_geometry(s.shape, cfg["tile_size"], cfg["overlap"]), sz, T, Zp) for s in stacks)
    vox = sum(s.numel() for s in stacks)
    batch = {"workload": "run_many over 40x2048^2, 40x2304^2 and 40x2800^2 uint16 + flats (25 + 25 + 36 tiles of "
                         "T 678) in shared launches", "ms": ms, "voxels_per_s": vox / (ms * 1e-3),
             "m1_roofline_frac": m1 / (ms * 1e-3) / HBM_PEAK, "spots": [int(o.shape[0]) for o in out]}
    log(f"uneven batch: {ms:.1f} ms for 3 stacks ({vox / ms / 1e6:.1f} G voxels/s, M1 {batch['m1_roofline_frac']:.3f})")
    del stacks, flats, out
    rng = np.random.default_rng(31)
    pairs = []
    for i, side in enumerate(sides):
        shift = (int(rng.integers(-3, 4)), int(rng.integers(-30, 31)), int(rng.integers(-30, 31)))
        feats = []
        a = time.perf_counter()
        for k, sh in enumerate(((0, 0, 0), shift)):
            nuc = nuclei_torch(7200 + i, (sz, side, side), int(2000 * side * side / 3000 ** 2), dev, shift=sh,
                               noise_seed=7300 + 2 * i + k)
            torch.cuda.synchronize()
            a0 = time.perf_counter()
            pl, mi = dapi_pipe(nuc)
            feats.append(types.SimpleNamespace(Xh_plus=pl.clone(), Xh_minus=mi.clone()))
            torch.cuda.synchronize()
            a += time.perf_counter() - a0
        b = time.perf_counter()
        f, tp, tm, n_p, n_m = DualAligner(feats[0], th=4).get_best_translation_pointsV2(feats[1])
        drift_ms = (time.perf_counter() - b) * 1e3
        ok = bool(np.array_equal(np.round(f).astype(int), -np.array(shift)))
        pairs.append({"side": side, "shift": list(shift), "tzxyf": [float(v) for v in f], "recovered": ok,
                      "N_plus": int(n_p), "N_minus": int(n_m), "drift_ms": drift_ms})
        log(f"drift {side}^2: shift {shift}, tzxyf {np.round(f, 3).tolist()}, recovered {ok}, {drift_ms:.1f} ms")
    torch.cuda.empty_cache()
    return {"batch": batch, "drift_pairs": pairs, "drift_all_recovered": all(p["recovered"] for p in pairs)}


def stream_requests(dev, cfg, n, seed=41):
    """Config-4 style requests: 3 x 40 x H x W uint16 (2 hyb channels + the DAPI channel, the reference's last
    channel), H = W uniform in {2048, 2304, 2800}, single or multi-PSF mode 50/50, FOV blocks of two
    acquisitions (the second DAPI shifted), host inputs in pinned memory.  Returns factories of FovRequests."""
    import torch
    from mermake_amd.synth import nuclei_torch, stack_torch
    rng = np.random.default_rng(seed)
    pool = {}
    for side in (2048, 2304, 2800):
        for k in range(2):
            chans, flat = [], None
            for c in range(2):
                s_, flat = stack_torch(8000 + 10 * side + 2 * k + c, (cfg["sz"], side, side), cfg["density"], dev,
                                       flat=True)
                chans.append(s_.cpu().pin_memory())
            sh = (0, 0, 0) if k == 0 else (1, 7, -11)
            nuc = nuclei_torch(8100 + side, (cfg["sz"], side, side), int(2000 * side * side / 3000 ** 2), dev,
                               shift=sh, noise_seed=8200 + 2 * side + k)
            pool[(side, k)] = (chans, flat.cpu().pin_memory(), nuc.cpu().pin_memory())
            torch.cuda.synchronize()
    plan = []
    for b in range(n // 2):
        side = int(rng.choice([2048, 2304, 2800]))
        for k in range(2):
            plan.append((b, side, k, "single" if rng.random() < 0.5 else "grid"))
    return pool, plan


def stream_leg(dev, cfg, pipes, dapi_pipe, args, rates):
    """The served stream: requests of ``stream_requests`` through one HybServer holding both PSF modes, the DAPI
    pipeline and the drift of each FOV block, at Poisson rates ``rates`` and saturated; against the plain form
    (one request at a time: the hyb channels one tile at a time with host syncs, the DAPI channel, the drift at
    the end of each block) with the same host-to-device copies."""
    import types
    import torch
    from mermake_amd.align import DualAligner
    from mermake_amd.pipeline import HybPipeline
    from mermake_amd.serve import FovRequest, HybServer, poisson_arrivals
    n = args.serve_requests
    pool, plan = stream_requests(dev, cfg, n)

    def make(arr):
        return [FovRequest(i, pool[(side, k)][0], pool[(side, k)][1], arrival=float(a), dapi=pool[(side, k)][2],
                           psf=mode, fov=b) for i, ((b, side, k, mode), a) in enumerate(zip(plan, arr))]

    srv = HybServer(pipes, batch_tiles={"single": 36, "grid": 50}, slots=2, dapi=dapi_pipe)
    srv.run(make(np.zeros(4))[:4])   # warm the slots and the DAPI buffers
    warm_copies(srv, list({id(t): t for chans, fl, nuc in pool.values() for t in list(chans) + [nuc]}.values()), dev)
    vox_per_req = [3 * cfg["sz"] * side * side for _, side, _, _ in plan]
    out = {"workload": f"{n} requests of 3 x 40 x H x W uint16 (2 hyb channels + DAPI), H = W uniform in "
                       "{2048, 2304, 2800}, single-PSF (tile 500) or 100-PSF keyed grid (tile 300) 50/50, FOV blocks "
                       "of 2 acquisitions with drift, host pinned inputs (H2D inside the loop)",
           "rates": {}}
    for rate in list(rates) + [0]:
        arr = poisson_arrivals(n, rate, seed=int(rate * 10) + 3) if rate > 0 else np.zeros(n)
        reqs = make(arr)
        with mark(f"stream_{'saturated' if rate == 0 else f'poisson_{rate:g}'}"):
            srv.run(reqs)
        st = dict(srv.stats)
        st["voxels_per_s"] = sum(vox_per_req) / st["wall_s"]
        st["drift_recovered"] = all(
            r.drift is not None and (r.fov is None or np.array_equal(
                np.round(r.drift[0]).astype(int), np.array((1, 7, -11)) * (1 if plan[r.rid][2] == 0 else 0)))
            for r in reqs)
        key = "saturated" if rate == 0 else f"poisson_{rate:g}"
        out["rates"][key] = st
        log(f"served stream {key}: {st['fovs_per_s']:.2f} FOV/s, p50 {st['p50_latency_s'] * 1e3:.0f} ms, "
            f"p99 {st['p99_latency_s'] * 1e3:.0f} ms, drift ok {st['drift_recovered']}")
    del srv
    # plain form: one request at a time
    plain = {k: HybPipeline(p.d, plain=True, blur_radius=p.blur_radius, threshold=p.threshold, delta=p.delta,
                            delta_fit=p.delta_fit, sigmaZ=p.sigmaZ, sigmaXY=p.sigmaXY) for k, p in pipes.items()}
    m = min(n, 8)
    torch.cuda.synchronize()
    a = time.perf_counter()
    feats = {}
    for b, side, k, mode in plan[:m]:
        chans, flat, nuc = pool[(side, k)]
        fl = flat.to(dev, non_blocking=True)
        for c in chans:
            plain[mode](c.to(dev, non_blocking=True), fl)
        pl, mi = dapi_pipe(nuc.to(dev, non_blocking=True))
        feats.setdefault(b, []).append(types.SimpleNamespace(Xh_plus=pl.clone(), Xh_minus=mi.clone()))
        if len(feats[b]) == 2:
            dual = DualAligner(feats[b][1], th=4)
            for f in feats[b]:
                dual.get_best_translation_pointsV2(f)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - a) / m
    out["plain_fovs_per_s"] = 1.0 / plain_s
    out["speedup_vs_plain"] = out["rates"]["saturated"]["fovs_per_s"] * plain_s
    log(f"served stream plain form: {1.0 / plain_s:.2f} FOV/s; saturated stream {out['speedup_vs_plain']:.2f}x")
    return out


def host_bytes(cfg):
    return cfg["channels"] * cfg["sz"] * cfg["sx"] * cfg["sy"] * 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-tile GPU form")
    ap.add_argument("--no-serve", action="store_true", help="skip the served (request loop) leg")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests of each served-stream run")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate of the served stream (FOV/s)")
    ap.add_argument("--serve-sweep", action="store_true",
                    help="run the served stream at Poisson rates 0.5, 1, 2, 3 and 4 FOV/s (about 4 minutes)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the multi-PSF grid, uneven-batch + drift and served-stream legs")
    ap.add_argument("--legs", default="multi_psf_grid,uneven_batch_drift,served_stream",
                    help="the legs to run (comma list of multi_psf_grid, uneven_batch_drift, served_stream)")
    ap.add_argument("--no-feed", action="store_true", help="skip the host feed leg (zarr decode, pinned H2D)")
    ap.add_argument("--no-dapi", action="store_true", help="skip the DAPI stack leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default schedule instead of timing every admitted one first")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph replay leg")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with torch.neg over 1 GiB of float32 (reads 1 GiB, writes 1 GiB): the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Tiles and FOVs are independent; see DESIGN.md for how the path would shard.", file=sys.stderr)
        sys.exit(2)
    import torch
    from mermake_amd import _lib as L
    from mermake_amd.deconvolver import tile_geometry
    from mermake_amd.pipeline import HybPipeline
    from mermake_amd.synth import gaussian_psf, stack_torch

    dev = torch.device("cuda", 0)
    cfg = CFG
    shape = (cfg["sz"], cfg["sx"], cfg["sy"])
    log("generating synthetic FOV", shape, "x", cfg["channels"])
    stacks, flat = [], None
    for c in range(cfg["channels"]):
        s, f = stack_torch(1000 + c, shape, cfg["density"], dev, flat=True)
        stacks.append(s)
        flat = f
    psf = gaussian_psf(cfg["psf_shape"], *cfg["psf_sigma"])
    torch.cuda.synchronize()
    log("building Wiener spectrum")
    hyb_kw = dict(blur_radius=cfg["blur_radius"], threshold=cfg["threshold"], delta=cfg["delta"],
                  delta_fit=cfg["delta_fit"], sigmaZ=cfg["sigmaZ"], sigmaXY=cfg["sigmaXY"])
    # the autotuner times every schedule the tile admits on this FOV's three channels (not cached:
    # the table is part of the record) and keeps the fastest
    tune = None if args.no_autotune else {"stacks": cfg["channels"], "reps": 2, "cache": False}
    if tune:
        log("autotuning the deconvolution schedule")
    pipe = HybPipeline.from_settings(psf, shape, tile_size=cfg["tile_size"], overlap=cfg["overlap"], beta=cfg["beta"],
                                     autotune=tune, **hyb_kw)
    d = pipe.d
    if d.tuning:
        log("schedule", d.tuning["xy"], d.tuning["z"], "tiles per launch", d.tuning["batch_tiles"] or "all",
            {k: round(v, 1) for k, v in d.tuning["ms_forms"].items()})
    T, Zp = d._plan.T, d._plan.Zp
    geom = tile_geometry(shape, cfg["tile_size"], cfg["overlap"], 1)

    def fov():  # the three channels' tiles batched into shared launches
        return pipe.run_many(stacks, [flat] * len(stacks))

    def submit():  # the same, enqueued on a side stream (two FOVs in flight)
        return pipe.submit(stacks, [flat] * len(stacks))

    for _ in range(args.warmup):
        out = fov()
    torch.cuda.synchronize()
    # per-pass timing with HIP events on the pipeline's stream (one profiled channel)
    L.call("mm_deconv_set_profiling", d._plan.h, 1)
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    pipe(stacks[0], flat)
    passes = (ctypes.c_float * 5)()
    L.call("mm_deconv_pass_times", d._plan.h, passes)
    L.call("mm_deconv_set_profiling", d._plan.h, 0)
    names = ["r2c", "x_fwd", "z_wiener", "x_inv", "c2r"]
    pt = dict(zip(names, [float(v) for v in passes]))
    pb = pass_bytes(geom, cfg["sz"], T, d._plan.wplanes, deconv_window(pipe, T))
    ev[0].record(st)
    pipe(stacks[0], flat)
    ev[1].record(st)
    torch.cuda.synchronize()
    stack_ms = ev[0].elapsed_time(ev[1])
    dom = max(pt, key=pt.get)
    achieved = pb[dom] / (pt[dom] * 1e-3)
    for _ in range(args.warmup):
        for p in [submit(), submit()]:
            out = p.result()
    torch.cuda.synchronize()
    log("timing", args.steps, "FOVs")
    # FOV steps as a stream: each FOV's kernel chain is enqueued while the previous one runs (two in flight), so
    # the host work between FOVs and the last small kernels of one FOV overlap the next FOV's passes; latency
    # is submit -> spots of that FOV
    lat = []
    with mark("fov_steps"):
        t0 = time.perf_counter()
        pend = []
        for _ in range(args.steps):
            pend.append((time.perf_counter(), submit()))
            if len(pend) == 2:
                a, p = pend.pop(0)
                out = p.result()
                lat.append(time.perf_counter() - a)
        while pend:
            a, p = pend.pop(0)
            out = p.result()
            lat.append(time.perf_counter() - a)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    # the same FOV steps one at a time (each waits for its spots before the next is enqueued)
    seq_lat = []
    with mark("fov_steps_sequential"):
        t1 = time.perf_counter()
        for _ in range(args.steps):
            a = time.perf_counter()
            seq_out = fov()
            torch.cuda.synchronize()
            seq_lat.append(time.perf_counter() - a)
        seq_ms = (time.perf_counter() - t1) / args.steps * 1e3
    same_seq = all(torch.equal(x, y) for x, y in zip(out, seq_out))
    del seq_out
    nspots = [int(o.shape[0]) for o in out]
    fov_vox = cfg["channels"] * cfg["sz"] * cfg["sx"] * cfg["sy"]
    value = fov_vox * args.steps / el
    ms = el / args.steps * 1e3

    m1 = m1_bytes_per_stack(geom, cfg["sz"], T, Zp)
    m1_frac = (m1 * cfg["channels"] / (ms * 1e-3)) / HBM_PEAK

    graph = None
    if not args.no_graph:
        # the same batched FOV step with each batch's kernel chain replayed as a hipGraph
        gp = HybPipeline(d, graph=True, **hyb_kw)
        gout = gp.run_many(stacks, [flat] * len(stacks))        # eager first batch, then the capture
        torch.cuda.synchronize()
        a = time.perf_counter()
        with mark("fov_steps_graph"):
            for _ in range(args.steps):
                gout = gp.run_many(stacks, [flat] * len(stacks))
            torch.cuda.synchronize()
        gms = (time.perf_counter() - a) / args.steps * 1e3
        same = all(torch.equal(x, y) for x, y in zip(gout, out))
        graph = {"ms_per_fov": gms, "eager_ms_per_fov": seq_ms, "graphs": len(gp._graphs), "equal_to_eager": same}
        del gp, gout
        torch.cuda.empty_cache()

    plain = None
    if not args.no_plain:
        pp = HybPipeline(d, plain=True, **hyb_kw)
        [pp(s, flat) for s in stacks]
        torch.cuda.synchronize()
        a = time.perf_counter()
        [pp(s, flat) for s in stacks]
        torch.cuda.synchronize()
        pms = (time.perf_counter() - a) * 1e3
        # the same with each channel copied from pinned host memory first, as the reference's
        # loop does with cp.asarray of each host channel
        hosts = [s.cpu().pin_memory() for s in stacks]
        torch.cuda.synchronize()
        a = time.perf_counter()
        [pp(h.to(dev, non_blocking=True), flat) for h in hosts]
        torch.cuda.synchronize()
        pms_h2d = (time.perf_counter() - a) * 1e3
        del hosts
        plain = {"ms_per_fov": pms, "fovs_per_s": 1e3 / pms, "batched_speedup": pms / ms,
                 "batched_speedup_sequential": pms / seq_ms,
                 "ms_per_fov_with_h2d": pms_h2d}

    served = None
    if not args.no_serve:
        served = serve_leg(pipe, stacks, flat, cfg, args, dev)
        if plain is not None:
            # voxel throughput of the saturated served stream (host inputs) over the plain form's
            # with the same host-to-device copies
            served["speedup_vs_plain_with_h2d"] = served["saturated"]["voxels_per_s"] / (
                fov_vox / (plain["ms_per_fov_with_h2d"] * 1e-3))

    dapi = None
    if not args.no_dapi:
        dapi = dapi_leg(stacks[0], flat, psf, cfg)

    legs = None
    if not args.no_legs:
        from mermake_amd.deconvolver import Deconvolver
        from mermake_amd.pipeline import DapiPipeline
        log("building the 100-PSF grid spectra (hyb and DAPI)")
        gpsfs = grid_psfs(10, 300)
        dapi_kw = dict(blur_radius=50, threshold=3.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5)
        gtune = None if args.no_autotune else {"reps": 2, "cache": False}
        grid = HybPipeline.from_settings(gpsfs, (cfg["sz"], 3000, 3000), tile_size=300, overlap=cfg["overlap"],
                                         beta=cfg["beta"], psf_mode="keyed", autotune=gtune, **hyb_kw)
        dapi_grid = DapiPipeline(Deconvolver(gpsfs, (cfg["sz"], 3000, 3000), tile_size=300, zpad=(cfg["sz"] - 1) // 2,
                                             overlap=cfg["overlap"], beta=0.01, psf_mode="keyed",
                                             autotune=gtune or False), **dapi_kw)
        dapi_one = DapiPipeline(Deconvolver(psf, shape, tile_size=cfg["tile_size"], zpad=(cfg["sz"] - 1) // 2,
                                            overlap=cfg["overlap"], beta=0.01), **dapi_kw)
        which = set(args.legs.split(","))
        legs = {}
        if "multi_psf_grid" in which:
            legs["multi_psf_grid"] = multi_psf_leg(dev, cfg, grid, dapi_grid)
        del dapi_grid
        torch.cuda.empty_cache()
        if "uneven_batch_drift" in which:
            legs["uneven_batch_drift"] = uneven_leg(dev, cfg, pipe, dapi_one)
        rates = [0.5, 1.0, 2.0, 3.0, 4.0] if args.serve_sweep else [args.serve_rate]
        if "served_stream" in which:
            legs["served_stream"] = stream_leg(dev, cfg, {"single": pipe, "grid": grid}, dapi_one, args, rates)
        del grid, dapi_one
        torch.cuda.empty_cache()

    feed = None
    if not args.no_feed:
        feed = feed_leg(dev, cfg)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, bounded sample)")
        cpu = cpu_baseline(stacks[0].cpu().numpy(), flat.cpu().numpy(), psf, cfg, len(geom))

    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    lat_sorted = sorted(lat)
    rec = {
        "metric": "FOV voxels/s (deconv->maxima->fit)", "value": value, "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded puncta + Poisson background, flat field)",
        "config": {"workload": "full hyb FOV: 3 channels x 40x2800x2800 uint16, tile 500 overlap 89 (36 tiles of "
                               "118x678x678 per channel), 1 PSF 40x64x64, beta 1e-4, blur 30, thr 3600, delta 1, "
                               "delta_fit 3", "fov_voxels": fov_vox, "spots_per_channel": nspots},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": measured_traffic(dom), "kernel": dom,
                     "kernel_ms": pt[dom], "kernel_algorithmic_bytes": pb[dom]},
        "m1_roofline_frac": m1_frac,
        "pass_ms": pt, "stack_ms": stack_ms,
        "schedule": {k: d.tuning[k] for k in ("xy", "z", "batch_tiles", "ms_forms", "ms_batch_tiles", "tune_s")}
        if d.tuning else d.schedule(),
        "graph": graph,
        "latency_ms": {"p50": 1e3 * lat_sorted[len(lat) // 2], "p99": 1e3 * lat_sorted[-1]},
        "sequential": {"ms_per_fov": seq_ms, "p50_ms": 1e3 * sorted(seq_lat)[len(seq_lat) // 2],
                       "equal_to_pipelined": same_seq,
                       "note": "the same FOV steps one at a time, each waiting for its spots"},
        "plain_gpu": plain,
        "served": served,
        "dapi": dapi,
        "legs": legs,
        "feed": feed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
