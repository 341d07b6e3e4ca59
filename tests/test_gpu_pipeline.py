"""GPU parity of the batched hyb / DAPI callers and of the drift."""
import types

import numpy as np
import pytest
import torch

from oracle import align as oalign
from oracle import deconv as odec
from oracle import maxima as omax

pytestmark = pytest.mark.gpu

HYB = dict(blur_radius=30, threshold=3600, delta=1, delta_fit=3, sigmaZ=1, sigmaXY=1.5)


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _match(a, b, tol):
    """Rows of b matched to rows of a by nearest (z,x,y); returns max |diff| of coordinates and the match rate."""
    if len(a) == 0 or len(b) == 0:
        return 0.0, float(len(a) == len(b))
    d = np.abs(a[:, None, :3] - b[None, :, :3]).max(-1)
    j = d.argmin(1)
    ok = d[np.arange(len(a)), j] < tol
    return float(d[np.arange(len(a)), j][ok].max(initial=0.0)), float(ok.mean())


@pytest.fixture(scope="module")
def small_hyb():
    """Config-0-like stack (scaled down): 20 x 300 x 300, 400 puncta, flat, tile 100, overlap 30."""
    from mermake_amd.synth import gaussian_psf, stack_np
    raw, flat = stack_np(11, (20, 300, 300), 400, flat=True)
    psf = gaussian_psf((21, 41, 41), 1.5, 1.2)
    return raw, flat, psf


def test_hyb_candidates_exact_given_same_volume(dev, small_hyb):
    """Given the identical deconvolved tile, maxima sets are bit-exact and fits within 1e-3 px."""
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.maxima import find_local_maxima, maxima_coords
    raw, flat, psf = small_hyb
    d = Deconvolver(psf, raw.shape, tile_size=100, zpad=19, overlap=30, beta=1e-4)
    n = 0
    for x, y, tile, rawv in d.tile_wise(T(raw, dev), flat_field=T(flat, dev), blur_radius=30):
        got = find_local_maxima(tile, threshold=3600, delta=1, delta_fit=3, raw=rawv).cpu().numpy()
        t_np, r_np = tile.cpu().numpy(), rawv.cpu().numpy()
        z, xx, yy = omax.local_maxima(t_np, np.float32(3600), 1)
        cand, _ = maxima_coords(tile, 3600.0, 1)
        assert np.array_equal(cand.cpu().numpy(), np.stack([z, xx, yy], 1).astype(np.int32)), (x, y)
        want = omax.find_local_maxima(t_np, threshold=3600, delta=1, delta_fit=3, raw=r_np)
        assert got.shape == want.shape, (x, y)
        if len(got):
            assert np.abs(got[:, :3] - want[:, :3]).max() < 1e-3
            assert np.array_equal(got[:, 7], want[:, 7].astype(np.float32))
        n += len(got)
    assert n > 50


def test_hyb_pipeline_vs_oracle(dev, small_hyb):
    """Batched hyb pipeline vs the float64 oracle of the reference loop (__main__.py:254-265)."""
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    pipe = HybPipeline.from_settings(psf, raw.shape, tile_size=100, overlap=30, beta=1e-4, **HYB)
    got = pipe(T(raw, dev), T(flat, dev)).cpu().numpy()
    W = odec.wiener(psf, raw.shape[0], 100, raw.shape[0] - 1, 30, 1e-4, mode="fp64")
    want = odec.hyb_spots(raw, W, flat, 100, 30, raw.shape[0] - 1, **HYB)
    assert abs(len(got) - len(want)) <= max(2, 0.01 * len(want)), (len(got), len(want))
    err, rate = _match(got, want, 0.5)
    assert rate > 0.99 and err < 1e-3, (err, rate)
    # plain (one tile at a time, host syncs) gives the same rows in the same order
    plain = HybPipeline(pipe.d, plain=True, **HYB)(T(raw, dev), T(flat, dev)).cpu().numpy()
    assert np.array_equal(plain, got)
    # determinism: a second run is bit-identical
    again = pipe(T(raw, dev), T(flat, dev)).cpu().numpy()
    assert np.array_equal(again, got)


@pytest.mark.parametrize("delta,delta_fit,ts,ov", [(1, 3, 100, 30), (2, 1, 100, 30), (1, 0, 100, 30),
                                                   (1, 3, 200, 13),   # T = 226: Rader passes, blocked spectra
                                                   (1, 3, 300, 89)])  # T = 478: Rader 239
def test_hyb_window_equals_whole_tiles(dev, small_hyb, delta, delta_fit, ts, ov):
    """Deconvolution, blur and maxima restricted to the margins the keep filter can use give the same kept rows,
    bit for bit, as whole tiles; the windowed candidate lists are the whole-tile lists restricted to the window."""
    from mermake_amd import _lib as L
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    hyb = dict(HYB, delta=delta, delta_fit=delta_fit, threshold=2000)
    win = HybPipeline.from_settings(psf, raw.shape, tile_size=ts, overlap=ov, beta=1e-4, spectral=False, **hyb)
    assert win.windows() is not None and not win.spectral_blur()
    full = HybPipeline(win.d, window=False, **hyb)
    a = win(T(raw, dev), T(flat, dev)).cpu().numpy()
    b = full(T(raw, dev), T(flat, dev)).cpu().numpy()
    assert len(a) > 50 and np.array_equal(a.view(np.uint32), b.view(np.uint32))  # delta_fit 0 fits NaN columns
    # candidates: whole-tile maxima restricted to the candidate window
    d = win.d
    inputs = _prepared(d, raw, flat, dev)   # the table points into these: keep them alive
    geom, table = full._table(*inputs)
    n = len(geom)
    (b0, b1), (c0, c1) = win.windows()
    Tt, sz = d._plan.T, d.tile_height
    res = d.run_tiles(table, n, n, 30).clone()
    ws_bytes = L.lib().mm_maxima_ws_bytes(sz, Tt, Tt, n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    cap = 8192
    out = []
    for name, extra in (("mm_tiles_maxima", ()), ("mm_tiles_maxima_window", (c0, c1))):
        coords = torch.zeros((n, cap, 3), dtype=torch.int32, device=dev)
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        L.call(name, d._plan.h, L.ptr(table), n, L.ptr(res), float(hyb["threshold"]), delta, *extra, L.ptr(coords),
               cap, L.ptr(counts), L.ptr(ws), ws_bytes, L.stream_ptr(dev))
        out.append((coords.cpu().numpy(), counts.cpu().numpy()))
    (cf, nf), (cw, nw) = out
    for t in range(n):
        f = cf[t, :nf[t]]
        keep = (f[:, 1] >= c0) & (f[:, 1] < c1) & (f[:, 2] >= c0) & (f[:, 2] < c1)
        assert np.array_equal(cw[t, :nw[t]], f[keep]), t


@pytest.mark.parametrize("ts,ov,zpad,shape", [(150, 38, 9, (20, 300, 280)),    # T 226: Rader 113
                                               (300, 89, 9, (20, 700, 640)),    # T 478: Rader 239
                                               (500, 89, 19, (40, 1100, 600))])  # T 678 (the DAPI tile class)
def test_deconv_tiles_stitch_equals_two_calls(dev, ts, ov, zpad, shape):
    """``mm_deconv_tiles_stitch`` (the inverse y transform storing each tile's owned part into the stitched volume)
    writes exactly what ``mm_deconv_tiles_window`` + ``mm_deconv_stitch`` write, into a strided output larger than
    the stack (the rest untouched), edge tiles included; ``apply`` uses it."""
    from mermake_amd import _lib as L
    from mermake_amd.deconvolver import Deconvolver, prepare_flat, prepare_stack, tile_table
    from mermake_amd.synth import gaussian_psf, stack_np
    raw, flat = stack_np(21, shape, 200, flat=True)
    d = Deconvolver(gaussian_psf((shape[0], 31, 31), 1.5, 1.2), shape, tile_size=ts, overlap=ov, zpad=zpad,
                    beta=1e-2)
    assert d.schedule()["xy"] == "rader"
    img = prepare_stack(T(raw, dev))
    fl = prepare_flat(T(flat, dev), img)
    geom = d.geometry(img.shape)
    table = tile_table(geom, img, fl)
    n, Tt, sz = len(geom), d._plan.T, shape[0]
    st = L.stream_ptr(dev)
    ops = d.box_spectra(50)
    spec = torch.empty(n * d._plan.spec_bytes, dtype=torch.uint8, device=dev)
    res = torch.empty((n, sz, Tt, Tt), dtype=torch.float32, device=dev)
    outs = []
    for fused in (False, True):
        big = torch.full((sz, shape[1] + 3, shape[2] + 5), float("nan"), dtype=torch.float32, device=dev)
        out = big[:, 1:1 + shape[1], 2:2 + shape[2]]
        if fused:
            L.call("mm_deconv_tiles_stitch", d._plan.h, L.ptr(table), n, L.ptr(ops), 1, n, L.ptr(spec), L.ptr(out),
                   out.stride(0), out.stride(1), shape[1], shape[2], st)
        else:
            L.call("mm_deconv_tiles_window", d._plan.h, L.ptr(table), n, L.ptr(ops), 1, n, L.ptr(spec), L.ptr(res),
                   ov, Tt - ov, st)
            L.call("mm_deconv_stitch", d._plan.h, L.ptr(table), n, L.ptr(res), L.ptr(out), out.stride(0),
                   out.stride(1), shape[1], shape[2], st)
        torch.cuda.synchronize(dev)
        outs.append(big.cpu().numpy())
    print("tiles_stitch", shape, ts, ov, Tt, n, flush=True)
    a, b = outs
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.isfinite(a[:, 1:1 + shape[1], 2:2 + shape[2]]).all() and np.isnan(a[:, 0]).all()
    # apply (fused) = the two calls
    got = d.apply(T(raw, dev), flat_field=T(flat, dev), blur_radius=50, spectral=True).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), a[:, 1:1 + shape[1], 2:2 + shape[2]].view(np.uint32))


@pytest.mark.parametrize("cap", [16384, 8])   # 8: every tile overflows the candidate cap, result() runs again
def test_submit_in_flight_equals_run_many(dev, small_hyb, cap):
    """``submit`` (two batches in flight on side streams, a third waiting for the first) gives ``run_many``'s
    rows bit for bit, for uneven stacks, host stacks and a cap every tile overflows."""
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    pipe = HybPipeline.from_settings(psf, raw.shape, tile_size=100, overlap=30, beta=1e-4, cap=cap, **HYB)
    stacks = [T(raw, dev), T(raw[:, :250, :280], dev), raw[:, 20:, 10:]]   # the last one from the host
    flats = [T(flat, dev), T(flat[:250, :280], dev), flat[20:, 10:]]
    want = [w.cpu().numpy() for w in pipe.run_many(stacks, flats)]
    assert all(len(w) > 20 for w in want)
    pend = [pipe.submit(stacks, flats) for _ in range(3)]
    for p in reversed(pend):
        got = [g.cpu().numpy() for g in p.result()]
        assert all(np.array_equal(g.view(np.uint32), w.view(np.uint32)) for g, w in zip(got, want))
    assert pend[0].result() is pend[0].result()


@pytest.mark.parametrize("delta,delta_fit,ts,ov", [(1, 3, 100, 30), (2, 1, 100, 30),
                                                   (1, 3, 150, 38),   # T = 226: Rader 113, blocked spectra
                                                   (1, 3, 300, 89)])  # T = 478: Rader 239
def test_hyb_spectral_blur_kept_rows_given_volume(dev, small_hyb, delta, delta_fit, ts, ov):
    """With the blur subtraction folded into the Wiener operator, the oracle's maxima, fit and keep filter on
    the pipeline's own volumes give exactly its kept rows, and those volumes equal the blur pass's over the
    blur window within float32 rounding (relative L2 <= 1e-5); the blur-pass pipeline's spots match."""
    from spot_rows import assert_kept_rows, oracle_tile_rows
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    hyb = dict(HYB, delta=delta, delta_fit=delta_fit, threshold=2000)
    sp = HybPipeline.from_settings(psf, raw.shape, tile_size=ts, overlap=ov, beta=1e-4, **hyb)
    assert sp.spectral_blur()
    ex = HybPipeline(sp.d, spectral=False, **hyb)
    a = sp(T(raw, dev), T(flat, dev)).cpu().numpy()
    b = ex(T(raw, dev), T(flat, dev)).cpu().numpy()
    va = sp.tile_volumes(T(raw, dev), T(flat, dev))
    vb = ex.tile_volumes(T(raw, dev), T(flat, dev))
    print("spectral blur", delta, delta_fit, sp.d._plan.T, len(a), len(b), flush=True)
    assert_kept_rows(a, oracle_tile_rows(va, ov, hyb), f"spectral blur T {sp.d._plan.T}")
    (lo, hi), _ = sp.windows()
    for (x, y, u, _), (_, _, v, _) in zip(va, vb):
        u, v = u[:, lo:hi, lo:hi], v[:, lo:hi, lo:hi]
        assert np.linalg.norm(u - v) <= 1e-5 * np.linalg.norm(v), (x, y)
    assert len(a) > 50 and abs(len(a) - len(b)) <= max(2, 0.01 * len(b)), (len(a), len(b))
    err, rate = _match(a, b, 0.5)
    assert rate > 0.99 and err < 1e-3, (err, rate)


@pytest.mark.parametrize("delta,blur_radius,thr,ts,ov,same_window", [
    (1, 30, 2000.0, 100, 30, False), (1, 30, -50.0, 100, 30, False),  # almost every voxel a candidate
    (3, 50, 500.0, 200, 13, False),   # T = 226, sorted-sphere test, blur half-width 25
    (5, 30, 1000.0, 100, 30, True),   # largest sphere; maxima window as wide as the blur window allows
    (2, 50, 800.0, 100, 30, False),
    (1, 30, float("nan"), 300, 89, False)])  # NaN threshold rejects nothing; T = 478
def test_blur_maxima_entry_equals_separate_calls(dev, small_hyb, delta, blur_radius, thr, ts, ov, same_window):
    """``mm_tiles_blur_maxima_window`` (threshold bits written by the blur, maxima from the bits) gives the
    blurred window and the candidate lists of ``mm_blur_sub_xy_window`` + ``mm_tiles_maxima_window``,
    bit for bit, edge tiles included."""
    from mermake_amd import _lib as L
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    hyb = dict(HYB, delta=delta, delta_fit=3, threshold=thr, blur_radius=blur_radius)
    win = HybPipeline.from_settings(psf, raw.shape, tile_size=ts, overlap=ov, beta=1e-4, **hyb)
    d = win.d
    inputs = _prepared(d, raw, flat, dev)   # the table points into these: keep them alive
    geom, table = win._table(*inputs)
    n = len(geom)
    (b0, b1), (c0, c1) = win.windows()
    if same_window:
        c0, c1 = b0 + delta, b1 - delta
    Tt, sz = d._plan.T, d.tile_height
    bd = blur_radius // 2
    print("blur+maxima entry", delta, blur_radius, thr, Tt, n, (b0, b1), (c0, c1), flush=True)
    res = d.run_tiles(table, n, n, None).clone()
    ws_bytes = L.lib().mm_maxima_ws_bytes(sz, Tt, Tt, n)
    cap = 1 << 15
    st = L.stream_ptr(dev)
    got = []
    for fused in (False, True):
        ws = torch.full((ws_bytes,), 0xA5, dtype=torch.uint8, device=dev)  # stale workspace contents
        out = torch.full_like(res, 7.0)
        coords = torch.zeros((n, cap, 3), dtype=torch.int32, device=dev)
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        if fused:
            L.call("mm_tiles_blur_maxima_window", d._plan.h, L.ptr(table), n, L.ptr(res), L.ptr(out), bd, b0, b1,
                   thr, delta, c0, c1, L.ptr(coords), cap, L.ptr(counts), L.ptr(ws), ws_bytes, st)
        else:
            L.call("mm_blur_sub_xy_window", L.ptr(res), L.ptr(out), None, n * sz, Tt, bd, b0, b1, st)
            L.call("mm_tiles_maxima_window", d._plan.h, L.ptr(table), n, L.ptr(out), thr, delta, c0, c1,
                   L.ptr(coords), cap, L.ptr(counts), L.ptr(ws), ws_bytes, st)
        torch.cuda.synchronize(dev)
        got.append((out[:, :, b0:b1, b0:b1].cpu().numpy(), coords.cpu().numpy(), counts.cpu().numpy()))
    (o1, k1, n1), (o2, k2, n2) = got
    assert np.array_equal(o1.view(np.uint32), o2.view(np.uint32))
    assert np.array_equal(n1, n2) and n1.sum() > 0
    for t in range(n):
        m = min(int(n1[t]), cap)
        assert np.array_equal(k1[t, :m], k2[t, :m]), t


@pytest.mark.parametrize("delta,thr,ts,ov,xy", [
    (1, 2000.0, 150, 38, "rader"),    # T = 226: Rader 113 C2R writes the threshold bits
    (1, -50.0, 150, 38, "rader"),     # almost every voxel a candidate
    (5, 1000.0, 300, 89, "rader"),    # T = 478: Rader 239; the largest sphere
    (2, float("nan"), 300, 89, "rader"),  # NaN threshold rejects nothing
    (1, 2000.0, 100, 30, None),       # T = 160: prime-first C2R, no bits (whole maxima pass)
    (1, 2000.0, 150, 38, "stockham")])
def test_deconv_maxima_entry_equals_separate_calls(dev, small_hyb, delta, thr, ts, ov, xy):
    """``mm_tiles_deconv_maxima_window`` (threshold bits written by the inverse y transform, maxima from the
    bits) gives the deconvolved window and the candidate lists of ``mm_deconv_tiles_window`` +
    ``mm_tiles_maxima_window``, bit for bit, edge tiles included, over a stale workspace."""
    from mermake_amd import _lib as L
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    hyb = dict(HYB, delta=delta, delta_fit=3, threshold=thr)
    win = HybPipeline.from_settings(psf, raw.shape, tile_size=ts, overlap=ov, beta=1e-4, **hyb)
    d = win.d
    if xy is not None:
        d.set_schedule(xy=xy)
    assert win.spectral_blur()
    inputs = _prepared(d, raw, flat, dev)   # the table points into these: keep them alive
    geom, table = win._table(*inputs)
    n = len(geom)
    (b0, b1), (c0, c1) = win.windows()
    Tt, sz = d._plan.T, d.tile_height
    ops = d.box_spectra(30)
    print("deconv+maxima entry", delta, thr, Tt, d.schedule()["xy"], n, (b0, b1), (c0, c1), flush=True)
    ws_bytes = L.lib().mm_maxima_ws_bytes(sz, Tt, Tt, n)
    spec = torch.empty(n * d._plan.spec_bytes, dtype=torch.uint8, device=dev)
    cap = 1 << 15
    st = L.stream_ptr(dev)
    got = []
    for fused in (False, True):
        ws = torch.full((ws_bytes,), 0xA5, dtype=torch.uint8, device=dev)  # stale workspace contents
        res = torch.full((n, sz, Tt, Tt), 7.0, dtype=torch.float32, device=dev)
        coords = torch.zeros((n, cap, 3), dtype=torch.int32, device=dev)
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        dargs = (d._plan.h, L.ptr(table), n, L.ptr(ops), ops.shape[0], n, L.ptr(spec), L.ptr(res), b0, b1)
        if fused:
            L.call("mm_tiles_deconv_maxima_window", *dargs, thr, delta, c0, c1, L.ptr(coords), cap, L.ptr(counts),
                   L.ptr(ws), ws_bytes, st)
        else:
            L.call("mm_deconv_tiles_window", *dargs, st)
            L.call("mm_tiles_maxima_window", d._plan.h, L.ptr(table), n, L.ptr(res), thr, delta, c0, c1,
                   L.ptr(coords), cap, L.ptr(counts), L.ptr(ws), ws_bytes, st)
        torch.cuda.synchronize(dev)
        got.append((res[:, :, b0:b1, b0:b1].cpu().numpy(), coords.cpu().numpy(), counts.cpu().numpy()))
    (o1, k1, n1), (o2, k2, n2) = got
    assert np.array_equal(o1.view(np.uint32), o2.view(np.uint32))
    assert np.array_equal(n1, n2) and n1.sum() > 0
    for t in range(n):
        m = min(int(n1[t]), cap)
        assert np.array_equal(k1[t, :m], k2[t, :m]), t


def test_deconv_maxima_entry_rejects_reads_outside_window(dev, small_hyb):
    """Maxima whose neighbour tests would read outside the computed window are refused."""
    from mermake_amd import _lib as L
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    win = HybPipeline.from_settings(psf, raw.shape, tile_size=150, overlap=38, beta=1e-4, **HYB)
    d = win.d
    inputs = _prepared(d, raw, flat, dev)   # the table points into these: keep them alive
    geom, table = win._table(*inputs)
    n = len(geom)
    (b0, b1), _ = win.windows()
    Tt, sz = d._plan.T, d.tile_height
    ops = d.box_spectra(30)
    ws_bytes = L.lib().mm_maxima_ws_bytes(sz, Tt, Tt, n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    spec = torch.empty(n * d._plan.spec_bytes, dtype=torch.uint8, device=dev)
    res = torch.empty((n, sz, Tt, Tt), dtype=torch.float32, device=dev)
    coords = torch.zeros((n, 16, 3), dtype=torch.int32, device=dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    with pytest.raises(L.MermakeLibraryError, match="reads outside"):
        L.call("mm_tiles_deconv_maxima_window", d._plan.h, L.ptr(table), n, L.ptr(ops), 1, n, L.ptr(spec),
               L.ptr(res), b0, b1, 2000.0, 2, b0 + 1, b1 - 2, L.ptr(coords), 16, L.ptr(counts), L.ptr(ws), ws_bytes,
               L.stream_ptr(dev))


@pytest.mark.parametrize("ts,ov", [(100, 30), (200, 13), (300, 89)])  # prime-first, Rader 113, Rader 239 loaders
def test_squared_u8_stack_equals_uint16_squares(dev, small_hyb, ts, ov):
    """A ``SquaredU8`` stack (uint8 values squared by the tile loader) gives the spots, the deconvolved
    tiles and the raw views of its uint16 squares, bit for bit."""
    from mermake_amd.deconvolver import SquaredU8
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf = small_hyb
    u8 = np.minimum(np.sqrt(raw), 255).astype(np.uint8)
    u8[0, :3, :5] = 255  # the largest square
    sq = u8.astype(np.uint16) ** 2
    hyb = dict(HYB, threshold=500)
    pipe = HybPipeline.from_settings(psf, raw.shape, tile_size=ts, overlap=ov, beta=1e-4, **hyb)
    a = pipe(SquaredU8.of(T(u8, dev)), T(flat, dev)).cpu().numpy()
    b = pipe(T(sq, dev), T(flat, dev)).cpu().numpy()
    assert len(a) > 20 and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    d = pipe.d
    # each generator's views live in buffers the next batch reuses: copy one side first
    ga = [(x, y, t.clone(), r.clone()) for x, y, t, r in d.tile_wise(SquaredU8.of(T(u8, dev)), flat_field=T(flat, dev))]
    gb = list(d.tile_wise(T(sq, dev), flat_field=T(flat, dev)))
    assert len(ga) == len(gb) > 0
    for (xa, ya, ta, ra), (xb, yb, tb, rb) in zip(ga, gb):
        assert (xa, ya) == (xb, yb)
        assert torch.equal(ta, tb) and torch.equal(ra, rb)


def _prepared(d, raw, flat, dev):
    from mermake_amd.deconvolver import prepare_flat, prepare_stack
    img = prepare_stack(T(raw, dev))
    return img, prepare_flat(T(flat, dev), img)


def test_dapi_pipeline_vs_oracle(dev):
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.pipeline import DapiPipeline
    from mermake_amd.synth import gaussian_psf, stack_np
    raw, flat = stack_np(12, (12, 160, 160), 60, amp=(3000, 20000), flat=True)
    psf = gaussian_psf((12, 31, 31), 1.5, 1.2)
    kw = dict(blur_radius=50, threshold=3.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5)
    d = Deconvolver(psf, raw.shape, tile_size=100, zpad=(raw.shape[0] - 1) // 2, overlap=20, beta=0.01)
    plus, minus = DapiPipeline(d, **kw)(T(raw, dev), T(flat, dev))
    W = odec.wiener(psf, raw.shape[0], 100, (raw.shape[0] - 1) // 2, 20, 0.01, mode="fp64")
    wp, wm = odec.dapi_spots(raw, W, flat, 100, 20, (raw.shape[0] - 1) // 2, **kw)
    for g, w in ((plus.cpu().numpy(), wp), (minus.cpu().numpy(), wm)):
        assert abs(len(g) - len(w)) <= max(2, 0.02 * len(w)), (len(g), len(w))
        err, rate = _match(g, w, 0.5)
        assert rate > 0.98 and err < 1e-3, (err, rate)


def test_dapi_pipeline_squared_u8_equals_uint16_squares(dev):
    """A ``SquaredU8`` DAPI stack gives the spots of its uint16 squares bit for bit: the deconvolution
    loaders and the fit's raw reads (``habs``, ``a``) both square the uint8 values."""
    from mermake_amd.deconvolver import Deconvolver, SquaredU8
    from mermake_amd.pipeline import DapiPipeline
    from mermake_amd.synth import gaussian_psf, stack_np
    raw, flat = stack_np(13, (12, 160, 160), 60, amp=(3000, 20000), flat=True)
    u8 = np.minimum(np.sqrt(raw), 255).astype(np.uint8)
    sq = u8.astype(np.uint16) ** 2
    psf = gaussian_psf((12, 31, 31), 1.5, 1.2)
    kw = dict(blur_radius=50, threshold=3.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5)
    d = Deconvolver(psf, raw.shape, tile_size=100, zpad=(raw.shape[0] - 1) // 2, overlap=20, beta=0.01)
    dp = DapiPipeline(d, **kw)
    pa, ma = [t.cpu().numpy() for t in dp(SquaredU8.of(T(u8, dev)), T(flat, dev))]
    pb, mb = [t.cpu().numpy() for t in dp(T(sq, dev), T(flat, dev))]
    assert len(pa) > 5
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert np.array_equal(ma.view(np.uint32), mb.view(np.uint32))
    # find_local_maxima with a SquaredU8 raw reads the squares too
    from mermake_amd.maxima import find_local_maxima
    img = torch.from_numpy(np.random.default_rng(3).random(u8.shape, dtype=np.float32)).to(dev)
    fa = find_local_maxima(img, threshold=0.9, delta=1, delta_fit=3, raw=SquaredU8.of(T(u8, dev))).cpu().numpy()
    fb = find_local_maxima(img, threshold=0.9, delta=1, delta_fit=3, raw=T(sq, dev)).cpu().numpy()
    assert len(fa) > 5 and np.array_equal(fa.view(np.uint32), fb.view(np.uint32))


def test_fftconvolve_vs_oracle(dev):
    from mermake_amd.utils import fftconvolve
    rng = np.random.default_rng(8)
    a = rng.random((5, 17, 23)).astype(np.float32)
    b = rng.random((3, 9, 4)).astype(np.float32)
    for mode in ("full", "same", "valid"):
        got = fftconvolve(T(a, dev), T(b, dev), mode=mode).cpu().numpy()
        want = oalign.fftconvolve(a, b, mode=mode)
        assert got.shape == want.shape
        assert np.abs(got - want).max() < 1e-4 * np.abs(want).max()


def _cloud(rng, n, lo=(0, 60, 60), hi=(40, 1400, 1400)):
    X = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    h = rng.uniform(4.5, 9, size=(n, 1)).astype(np.float32)
    return np.concatenate([X, np.zeros((n, 4), np.float32), h], 1)


@pytest.mark.parametrize("shift", [(2, -15, 25), (-3, 30, -30), (0, 0, 0), (1, 7, -12)])
def test_drift_recovers_integer_shift(dev, shift):
    from mermake_amd.align import Aligner
    rng = np.random.default_rng(9)
    ref = _cloud(rng, 800)
    mov = ref.copy()
    mov[:, :3] += np.array(shift, np.float32)
    mov[:, :3] += rng.normal(0, 0.2, size=(len(mov), 3)).astype(np.float32)
    al = Aligner(T(ref[:, :3], dev))
    t, n = al.get_best_translation_points(T(mov[:, :3], dev), return_counts=True)
    assert np.array_equal(np.round(t).astype(int), np.array(shift)), t
    ot, on = oalign.Aligner(ref[:, :3]).get_best_translation_points(mov[:, :3], return_counts=True)
    assert on == n and np.abs(ot - t).max() < 1e-4
    assert np.array_equal(al.correlation_peak(T(mov[:, :3], dev)), oalign.Aligner(ref[:, :3]).correlation_peak(
        mov[:, :3]))


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (37, 15), (300, 257), (1000, 700)])
def test_nn_round_nearest_and_lowest_index_on_ties(dev, n, m):
    """mm_nn_round's nearest reference point of every point, against a float64 brute force: the same distance
    and, among equally near points (duplicated reference points, lattice points at equal distances), the
    lowest index, as scipy's cKDTree query of the reference (align.py:61-77) returns for a unique nearest."""
    from mermake_amd import _lib as L
    rng = np.random.default_rng(n + m)
    R = rng.integers(0, 6, (m, 3)).astype(np.float32)      # a small lattice: many equal distances
    if m > 2:
        R[m // 2] = R[1]                                      # a duplicate point
    X = (rng.integers(0, 6, (n, 3)) + rng.choice([0.0, 0.5], (n, 3))).astype(np.float32)
    t = np.array([0.25, -0.5, 0.0])
    tx, tr = T(X, dev), T(R, dev)
    td = torch.tensor(t, dtype=torch.float64, device=dev)
    ds = torch.empty(n, dtype=torch.float64, device=dev)
    ix = torch.empty(n, dtype=torch.int32, device=dev)
    npts = torch.zeros(1, dtype=torch.int32, device=dev)
    L.call("mm_nn_round", L.ptr(tx), n, 3, L.ptr(tr), m, 3, L.ptr(td), 1e9, L.ptr(ds), L.ptr(ix), L.ptr(npts),
           L.stream_ptr(dev))
    d2 = (((X.astype(np.float64) - t)[:, None, :] - R.astype(np.float64)[None, :, :]) ** 2).sum(-1)
    want_i = d2.argmin(1)                                      # first index of the minimum
    got_i = ix.cpu().numpy()
    assert np.array_equal(got_i, want_i)
    assert np.allclose(ds.cpu().numpy(), np.sqrt(d2[np.arange(n), want_i]), rtol=0, atol=1e-12)
    assert int(npts.item()) == n


def test_drift_fixture_known_answer(dev):
    """Reference drift pickle: N+ = 621, N- = 261, zero shift (y > 4 reference-threshold quirk)."""
    from fixtures import dapi_features, drift_known_answer
    from mermake_amd.align import DualAligner, drift
    xp, xm = dapi_features()
    ka = drift_known_answer()
    img = types.SimpleNamespace(Xh_plus=T(xp, dev), Xh_minus=T(xm, dev), path="/data/H1_TEST_set1/Conv_zscan__020.zarr")
    dual = DualAligner(img, th=4)
    f, tp, tm, n_p, n_m = dual.get_best_translation_pointsV2(img)
    assert (n_p, n_m) == (ka["N_plus"], ka["N_minus"])
    assert np.allclose(f, ka["tzxyf"]) and np.allclose(tp, 0) and np.allclose(tm, 0)

    class Block(list):
        def ifov(self):
            return 20

        def iset(self):
            return 1

        def drift_file(self):
            return "/nonexistent/drift_20--_set1.pkl"

    res, path = drift(Block([img]))
    drifts, files, ifov, refpath = res
    assert len(drifts) == 1 and drifts[0][3] == 621 and drifts[0][4] == 261 and ifov == 20
