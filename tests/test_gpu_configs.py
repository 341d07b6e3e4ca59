"""GPU parity at the tile classes of the five benchmark workloads.

Every test runs the HIP path through the C ABI at a workload's own geometry (stack depth, tile size,
overlap, zpad, PSF set, flat) and checks it against the float64 oracle:

* deconvolution, per tile: relative L2 <= 1e-5 against the oracle's float64 ``tile_wise``;
* the volume the maxima read (deconvolved tile minus its blur, the blur folded into the Wiener operator):
  relative L2 <= 1e-5 against the oracle's float64 ``tile_wise`` with the blur, over the blur window;
* spots, given the GPU's own volumes (``HybPipeline.tile_volumes``): the oracle's maxima, fit and keep
  filter (``maxima.py:17-98``, ``__main__.py:254-265``) run on those volumes must give exactly the kept rows of
  ``HybPipeline`` (same rows in the same order; fitted coordinates within 1e-3 px; background, habs and h
  equal; correlations within 1e-4).  A row whose float64 fitted x or y lies within 1e-3 px of a keep
  boundary may fall on either side;
* end to end against the float64 oracle's own deconvolution: the spot count within 0.2 % and at least
  99.5 % of the spots matched within 1e-3 px (the remaining ones are candidates whose value or neighbour
  order changes with the ~1e-6 relative deconvolution differences).

Workloads (sizes reduced only where the oracle would take minutes; the tile class is the workload's):
  plain method     20x512^2 whole (4 tiles of T 678, Zp 58, 21x41x41 PSF, no flat)
  full hyb FOV     40x1200^2 with tile 500 / overlap 89, zpad 39, quadratic flat, 40x64x64 PSF:
                   corner, edge, interior and far-corner tiles of T 678 (Rader 113, column-blocked
                   spectra, length-80 z pass)
  multi-PSF grid   40x900^2 at tile 300 with a 3x3 dict of distinct PSFs (T 478, one PSF per tile), and
                   the DAPI pass (zpad 19, Zp 78, beta 0.01, blur 50, delta 5) on nuclei
  uneven batch     40x{2048, 2304, 2800}^2 FOVs in one batch, bit-equal to each alone; DAPI pairs shifted
                   by seeded integer offsets, recovered exactly by the drift
"""
import types

import numpy as np
import pytest
import torch

from oracle import deconv as odec
from oracle import maxima as omax

pytestmark = pytest.mark.gpu

HYB = dict(blur_radius=30, threshold=3600, delta=1, delta_fit=3, sigmaZ=1, sigmaXY=1.5)
DAPI = dict(blur_radius=50, threshold=3.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5)


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def log(*a):
    print(*a, flush=True)


# ------------------------------------------------------------------ comparison helpers
from spot_rows import BOUNDARY_TOL, _close, assert_kept_rows, oracle_tile_rows  # noqa: E402,F401


def gpu_tiles(d, stack, flat, blur_radius, keep=None):
    """Host copies (x, y, tile, raw) of the GPU's ``tile_wise`` views (valid only until the next batch)."""
    out = []
    for x, y, t, r in d.tile_wise(stack, flat_field=flat, blur_radius=blur_radius):
        if keep is None or (x, y) in keep:
            out.append((x, y, t.cpu().numpy(), r.cpu().numpy()))
    return out


def assert_deconv_parity(d, stack_d, flat_d, stack, flat, W, ts, ov, zpad, only, blur_radius=None, what=""):
    mine = {(x, y): t for x, y, t, _ in gpu_tiles(d, stack_d, flat_d, blur_radius, only)}
    n = 0
    for x, y, t, _ in odec.tile_wise(stack, W, flat, blur_radius, ts, ov, zpad, only=only):
        g = mine[(x, y)]
        assert g.shape == t.shape, (what, x, y, g.shape, t.shape)
        e = float(np.linalg.norm(g - t) / np.linalg.norm(t))
        log(f"{what} tile ({x}, {y}) {g.shape}: deconvolution relative L2 {e:.2e}")
        assert e <= 1e-5, (what, x, y, e)
        n += 1
    assert n == len(only)


def assert_volume_parity(pipe, stack_d, flat_d, stack, flat, W, ts, ov, zpad, only, what=""):
    """The volume ``pipe``'s maxima and fit read (with the blur folded into the Wiener operator) against the
    float64 oracle's ``tile_wise`` with the blur (``deconvolver.py:155-163``), over the blur window."""
    assert pipe.spectral_blur(), what
    (lo, hi), _ = pipe.windows()
    mine = {(x, y): v for x, y, v, _ in pipe.tile_volumes(stack_d, flat_d)}
    n = 0
    for x, y, t, _ in odec.tile_wise(stack, W, flat, pipe.blur_radius, ts, ov, zpad, only=only):
        g, w = mine[(x, y)][:, lo:hi, lo:hi], t[:, lo:hi, lo:hi]
        assert g.shape == w.shape, (what, x, y, g.shape, w.shape)
        e = float(np.linalg.norm(g - w) / np.linalg.norm(w))
        log(f"{what} tile ({x}, {y}) window [{lo}, {hi}): deconvolution minus blur, relative L2 {e:.2e}")
        assert e <= 1e-5, (what, x, y, e)
        n += 1
    assert n == len(only)


def assert_end_to_end(got, want, what):
    """HIP spots (own deconvolution) vs the float64 oracle's spots (float64 deconvolution)."""
    assert abs(len(got) - len(want)) <= max(2, 0.002 * len(want)), (what, len(got), len(want))
    d = np.abs(got[:, None, :3].astype(np.float64) - want[None, :, :3]).max(-1) if len(got) and len(want) else None
    if d is None:
        assert len(got) == len(want)
        return
    j = d.argmin(1)
    best = d[np.arange(len(got)), j]
    ok = best <= 1e-3
    log(f"{what}: {len(got)} spots vs {len(want)}, {ok.mean():.4f} matched, worst matched {best[ok].max():.1e} px")
    assert ok.mean() >= 0.995, (what, ok.mean())


# ------------------------------------------------------------------ plain method, small (whole workload)
@pytest.fixture(scope="module")
def plain_small():
    from mermake_amd.synth import gaussian_psf, stack_np
    raw, _ = stack_np(1000, (20, 512, 512), 2000)
    psf = gaussian_psf((21, 41, 41), 1.5, 1.2)
    W = odec.wiener(psf, 20, 500, 19, 89, 1e-4, mode="fp64")
    return raw, psf, W


def test_plain_small_deconvolution(dev, plain_small):
    from mermake_amd.deconvolver import Deconvolver
    raw, psf, W = plain_small
    d = Deconvolver(psf, raw.shape, tile_size=500, zpad=19, overlap=89, beta=1e-4)
    assert (d._plan.T, d._plan.Zp) == (678, 58)
    only = {(0, 0), (0, 500), (500, 0), (500, 500)}
    assert_deconv_parity(d, T(raw, dev), None, raw, None, W, 500, 89, 19, only, what="plain 20x512^2")


def test_plain_small_spots(dev, plain_small):
    from mermake_amd.pipeline import HybPipeline
    raw, psf, W = plain_small
    pipe = HybPipeline.from_settings(psf, raw.shape, tile_size=500, overlap=89, beta=1e-4, **HYB)
    got = pipe(T(raw, dev)).cpu().numpy()
    tiles = pipe.tile_volumes(T(raw, dev))
    assert len(tiles) == 4
    assert_kept_rows(got, oracle_tile_rows(tiles, 89, HYB), "plain 20x512^2")
    assert_volume_parity(pipe, T(raw, dev), None, raw, None, W, 500, 89, 19, {(0, 0), (500, 500)},
                         what="plain 20x512^2")
    want = odec.hyb_spots(raw, W, None, 500, 89, 19, **HYB)
    assert_end_to_end(got, want, "plain 20x512^2 end to end")


# ------------------------------------------------------------------ full hyb FOV tile class
@pytest.fixture(scope="module")
def hyb_fov():
    from mermake_amd.synth import gaussian_psf, stack_np
    raw, flat = stack_np(1001, (40, 1200, 1200), int(40 * 1200 * 1200 / 20 ** 3), flat=True)
    psf = gaussian_psf((40, 64, 64), 1.5, 1.2)
    W = odec.wiener(psf, 40, 500, 39, 89, 1e-4, mode="fp64")
    return raw, flat, psf, W


def test_hyb_fov_tile_class_deconvolution(dev, hyb_fov):
    from mermake_amd.deconvolver import Deconvolver
    raw, flat, psf, W = hyb_fov
    d = Deconvolver(psf, raw.shape, tile_size=500, zpad=39, overlap=89, beta=1e-4)
    assert (d._plan.T, d._plan.Zp, d._plan.blk, d._plan.wplanes) == (678, 118, 6, 240)
    only = {(0, 0), (0, 500), (500, 500), (1000, 1000)}
    assert_deconv_parity(d, T(raw, dev), T(flat, dev), raw, flat, W, 500, 89, 39, only, what="hyb 40x1200^2")


def test_hyb_fov_tile_class_spots(dev, hyb_fov):
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psf, W = hyb_fov
    pipe = HybPipeline.from_settings(psf, raw.shape, tile_size=500, overlap=89, beta=1e-4, **HYB)
    assert pipe.windows() is not None
    got = pipe(T(raw, dev), T(flat, dev)).cpu().numpy()
    tiles = pipe.tile_volumes(T(raw, dev), T(flat, dev))
    assert len(tiles) == 9
    assert_kept_rows(got, oracle_tile_rows(tiles, 89, HYB), "hyb 40x1200^2")
    assert_volume_parity(pipe, T(raw, dev), T(flat, dev), raw, flat, W, 500, 89, 39, {(0, 500), (1000, 1000)},
                         what="hyb 40x1200^2")
    # end to end on the corner, edge, interior and far-corner tiles' own spots
    only = {(0, 0), (0, 500), (500, 500), (1000, 1000)}
    want = odec.hyb_spots(raw, W, flat, 500, 89, 39, tile_filter=only, **HYB)
    mine = _rows_of_tiles(got, only, 500, 89, raw.shape)
    assert_end_to_end(mine, want, "hyb 40x1200^2 end to end (4 tiles)")


def _rows_of_tiles(rows, only, ts, ov, shape):
    """Stack rows whose fitted x, y fall in the owned region of one of the tiles ``only``."""
    keep = np.zeros(len(rows), bool)
    for x, y in only:
        keep |= (rows[:, 1] >= x) & (rows[:, 1] < min(x + ts, shape[1])) & (rows[:, 2] >= y) & (
            rows[:, 2] < min(y + ts, shape[2]))
    return rows[keep]


# ------------------------------------------------------------------ multi-PSF grid
def _grid_psfs(n=3, shape=(40, 64, 64), ts=300):
    """n x n PSFs keyed (0, x, y) by tile centre in x-major order, widths varying smoothly across the grid
    (sigma_xy 1.0 -> 1.6, sigma_z 1.2 -> 2.0)."""
    from mermake_amd.synth import gaussian_psf
    out = {}
    for i in range(n):
        for j in range(n):
            t = (i * n + j) / max(1, n * n - 1)
            out[(0, i * ts + ts // 2, j * ts + ts // 2)] = gaussian_psf(shape, 1.2 + 0.8 * t, 1.0 + 0.6 * t)
    return out


@pytest.fixture(scope="module")
def multi_psf():
    from mermake_amd.synth import stack_np
    raw, flat = stack_np(1002, (40, 900, 900), int(40 * 900 * 900 / 20 ** 3), flat=True)
    psfs = _grid_psfs()
    W = odec.wiener(psfs, 40, 300, 39, 89, 1e-4, mode="fp64")
    return raw, flat, psfs, W


def test_multi_psf_grid_deconvolution(dev, multi_psf):
    from mermake_amd.deconvolver import Deconvolver
    raw, flat, psfs, W = multi_psf
    d = Deconvolver(psfs, raw.shape, tile_size=300, zpad=39, overlap=89, beta=1e-4)
    assert (d._plan.T, d._plan.Zp, d.psf_fft.shape[0]) == (478, 118, 9)
    assert [t["psf"] for t in d.geometry(raw.shape)] == list(range(9))
    only = {(0, 0), (0, 300), (300, 300), (600, 600), (600, 0)}
    assert_deconv_parity(d, T(raw, dev), T(flat, dev), raw, flat, W, 300, 89, 39, only, what="multi-PSF 40x900^2")


def test_multi_psf_grid_spots(dev, multi_psf):
    from mermake_amd.pipeline import HybPipeline
    raw, flat, psfs, W = multi_psf
    pipe = HybPipeline.from_settings(psfs, raw.shape, tile_size=300, overlap=89, beta=1e-4, **HYB)
    got = pipe(T(raw, dev), T(flat, dev)).cpu().numpy()
    tiles = pipe.tile_volumes(T(raw, dev), T(flat, dev))
    assert len(tiles) == 9
    assert_kept_rows(got, oracle_tile_rows(tiles, 89, HYB), "multi-PSF 40x900^2")
    assert_volume_parity(pipe, T(raw, dev), T(flat, dev), raw, flat, W, 300, 89, 39, {(300, 300), (600, 0)},
                         what="multi-PSF 40x900^2")


@pytest.fixture(scope="module")
def dapi_stack(dev):
    from mermake_amd.synth import gaussian_psf, nuclei_torch
    raw = nuclei_torch(1003, (40, 900, 900), 180, dev).cpu().numpy()
    psf = gaussian_psf((40, 64, 64), 1.5, 1.2)
    return raw, psf


def test_multi_psf_grid_dapi_deconvolution(dev, dapi_stack):
    """The DAPI pass at zpad (sz - 1) // 2 = 19 (Zp 78), beta 0.01, with its blur of radius 50."""
    from mermake_amd.deconvolver import Deconvolver
    raw, psf = dapi_stack
    d = Deconvolver(psf, raw.shape, tile_size=300, zpad=19, overlap=89, beta=0.01)
    assert (d._plan.T, d._plan.Zp) == (478, 78)
    W = odec.wiener(psf, 40, 300, 19, 89, 0.01, mode="fp64")
    only = {(0, 0), (300, 300), (600, 0)}
    assert_deconv_parity(d, T(raw, dev), None, raw, None, W, 300, 89, 19, only, blur_radius=50,
                         what="DAPI 40x900^2")


def test_multi_psf_grid_dapi_own_psf_deconvolution(dev, dapi_stack):
    """The DAPI pass with the 3x3 PSF grid (one PSF per tile, zpad 19, Zp 78): the z FFT form of the z pass
    (frequency-domain W, 78 planes; its length-78 transforms in registers), against the float64 oracle."""
    from mermake_amd.deconvolver import Deconvolver
    raw, _ = dapi_stack
    psfs = _grid_psfs()
    d = Deconvolver(psfs, raw.shape, tile_size=300, zpad=19, overlap=89, beta=0.01)
    assert (d._plan.T, d._plan.Zp, d._plan.wplanes, d.psf_fft.shape[0]) == (478, 78, 78, 9)
    assert d.schedule()["z"] == "fft"
    W = odec.wiener(psfs, 40, 300, 19, 89, 0.01, mode="fp64")
    only = {(0, 0), (300, 300), (600, 0), (600, 600)}
    assert_deconv_parity(d, T(raw, dev), None, raw, None, W, 300, 89, 19, only, blur_radius=50,
                         what="DAPI 40x900^2 3x3 PSF grid")


def test_dapi_spectral_apply_vs_oracle(dev, dapi_stack):
    """DapiPipeline's volume before normalisation: ``apply`` with the blur of radius 50 folded into the
    Wiener operator (every box of the stitched window stays inside its tile) against the float64 oracle's
    ``apply`` with the blur pass, and against the HIP blur pass itself."""
    from mermake_amd.deconvolver import Deconvolver
    raw, psf = dapi_stack
    d = Deconvolver(psf, raw.shape, tile_size=300, zpad=19, overlap=89, beta=0.01)
    T_, ov = d._plan.T, d.overlap
    assert d.spectral_blur(50, (ov, T_ - ov))
    W = odec.wiener(psf, 40, 300, 19, 89, 0.01, mode="fp64")
    got = d.apply(T(raw, dev), blur_radius=50, spectral=True).cpu().numpy()
    exact = d.apply(T(raw, dev), blur_radius=50).cpu().numpy()
    want = odec.apply(raw, W, None, 50, 300, 89, 19)
    e = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    e2 = float(np.linalg.norm(exact - want) / np.linalg.norm(want))
    log(f"DAPI 40x900^2 apply, blur on the operator: relative L2 {e:.2e} (blur pass {e2:.2e})")
    assert e <= 1e-5 and e2 <= 1e-5, (e, e2)


def test_multi_psf_grid_dapi_spots_given_volume(dev, dapi_stack):
    """DapiPipeline's + and - spots equal the oracle's maxima and fits on the GPU's own normalised volume
    (``__main__.py:217-225``), row for row."""
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.pipeline import DapiPipeline
    raw, psf = dapi_stack
    d = Deconvolver(psf, raw.shape, tile_size=300, zpad=19, overlap=89, beta=0.01)
    dp = DapiPipeline(d, **DAPI)
    plus, minus = (t.cpu().numpy() for t in dp(T(raw, dev)))
    buf = dp._buffer.cpu().numpy()
    rms = np.sqrt(np.sum(buf.astype(np.float64) ** 2) / buf.size)
    assert abs(rms - 1.0) < 1e-5, rms
    kw = dict(DAPI)
    kw.pop("blur_radius")
    for sign, got in ((1, plus), (-1, minus)):
        want = omax.find_local_maxima(sign * buf, raw=raw.astype(np.float32), dtype=np.float64, **kw)
        log(f"DAPI 40x900^2 sign {sign:+d}: {len(got)} spots (oracle {len(want)})")
        assert len(got) == len(want) and len(got) > 0, (sign, len(got), len(want))
        for g, w in zip(got, want):
            assert _close(g, w), (sign, g, w)


# ------------------------------------------------------------------ uneven batch + drift
def test_uneven_fov_batch_equals_alone(dev):
    """Tiles of 40x2048^2, 40x2304^2 and 40x2800^2 FOVs (interior and edge tiles of different sizes) in shared
    launches give each FOV's spots of it alone, bit for bit."""
    from mermake_amd.pipeline import HybPipeline
    from mermake_amd.synth import gaussian_psf, stack_torch
    psf = gaussian_psf((40, 64, 64), 1.5, 1.2)
    stacks, flats = [], []
    for i, side in enumerate((2048, 2304, 2800)):
        s, f = stack_torch(3000 + i, (40, side, side), 1.0 / 20 ** 3, dev, flat=True)
        stacks.append(s)
        flats.append(f)
    pipe = HybPipeline.from_settings(psf, (40, 2800, 2800), tile_size=500, overlap=89, beta=1e-4, **HYB)
    alone = [pipe(s, f).cpu().numpy() for s, f in zip(stacks, flats)]
    together = pipe.run_many(stacks, flats)
    assert [len(a) for a in alone] == [len(t) for t in together]
    for i, (a, t) in enumerate(zip(alone, together)):
        log(f"uneven batch FOV {i}: {len(a)} spots")
        assert len(a) > 1000 and np.array_equal(a.view(np.uint32), t.cpu().numpy().view(np.uint32)), i


@pytest.mark.parametrize("shift", [(2, -17, 23), (-3, 30, -29)])
def test_dapi_pair_drift_recovers_integer_shift(dev, shift):
    """A DAPI stack and the same scene shifted by a seeded integer offset (with independent noise) through
    DapiPipeline and DualAligner (``align.py:131-156``, fed as ``__main__.py:221-225``): the combined drift and
    the plus side recover the shift exactly (the minus side too when it has enough matched points)."""
    from mermake_amd.align import DualAligner
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.pipeline import DapiPipeline
    from mermake_amd.synth import gaussian_psf, nuclei_torch
    shape = (40, 2048, 2048)
    psf = gaussian_psf((40, 64, 64), 1.5, 1.2)
    d = Deconvolver(psf, shape, tile_size=500, zpad=19, overlap=89, beta=0.01)
    dp = DapiPipeline(d, **DAPI)
    feats = []
    for k, sh in enumerate(((0, 0, 0), shift)):
        stack = nuclei_torch(4000, shape, 930, dev, shift=sh, noise_seed=4100 + k)
        plus, minus = dp(stack)
        feats.append(types.SimpleNamespace(Xh_plus=plus.clone(), Xh_minus=minus.clone()))
    log(f"drift {shift}: features + {len(feats[0].Xh_plus)}/{len(feats[1].Xh_plus)}, "
        f"- {len(feats[0].Xh_minus)}/{len(feats[1].Xh_minus)}")
    f, tp, tm, n_p, n_m = DualAligner(feats[0], th=4).get_best_translation_pointsV2(feats[1])
    log(f"drift {shift}: tzxyf {f}, plus {tp} ({n_p}), minus {tm} ({n_m})")
    assert n_p > 1000
    assert np.array_equal(np.round(tp).astype(int), np.array(shift)), tp
    if n_m >= 50:      # the minus side (maxima of -buffer with h > 4) is weak on this scene
        assert np.array_equal(np.round(tm).astype(int), np.array(shift)), tm
    assert np.array_equal(np.round(f).astype(int), -np.array(shift)), f


# ------------------------------------------------------------------ served stream
def test_served_stream_full_sizes_equal_alone(dev):
    """Served stream at full FOV sizes: six requests of 2 hyb channels + 1 DAPI channel, 40x{2048, 2304,
    2800}^2, each FOV acquired twice (the second DAPI shifted by an integer offset), the single-PSF and the
    keyed PSF-grid modes alternating, arriving as a Poisson stream.  Every request's hyb spots, DAPI features
    and block drift equal those of the same pipelines alone, and each block's drift recovers its shift."""
    from mermake_amd.align import DualAligner
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.pipeline import DapiPipeline, HybPipeline
    from mermake_amd.serve import FovRequest, HybServer, poisson_arrivals
    from mermake_amd.synth import gaussian_psf, nuclei_torch, stack_torch
    psf = gaussian_psf((40, 64, 64), 1.5, 1.2)
    grid_psfs = {}
    for i in range(4):
        for j in range(4):
            t = (i * 4 + j) / 15
            grid_psfs[(0, 350 + 700 * i, 350 + 700 * j)] = gaussian_psf((40, 64, 64), 1.2 + 0.8 * t, 1.0 + 0.6 * t)
    single = HybPipeline.from_settings(psf, (40, 2800, 2800), tile_size=500, overlap=89, beta=1e-4, **HYB)
    grid = HybPipeline.from_settings(grid_psfs, (40, 2800, 2800), tile_size=300, overlap=89, beta=1e-4,
                                     psf_mode="keyed", **HYB)
    dapi = DapiPipeline(Deconvolver(psf, (40, 2800, 2800), tile_size=500, zpad=19, overlap=89, beta=0.01), **DAPI)
    plan = [("A", 2048, "single", (0, 0, 0)), ("A", 2048, "grid", (1, -12, 9)),
            ("B", 2304, "grid", (0, 0, 0)), ("B", 2304, "single", (-2, 21, 14)),
            ("C", 2800, "single", (0, 0, 0)), ("C", 2800, "grid", (2, 8, -25))]
    blobs = {2048: 930, 2304: 1180, 2800: 1740}
    arrivals = poisson_arrivals(len(plan), 8.0, seed=5)
    reqs, want = [], []
    for i, (fov, side, mode, sh) in enumerate(plan):
        chans, flat = [], None
        for c in range(2):
            s, flat = stack_torch(5000 + 10 * i + c, (40, side, side), 1.0 / 20 ** 3, dev, flat=True)
            chans.append(s)
        da = nuclei_torch(6000 + ord(fov), (40, side, side), blobs[side], dev, shift=sh, noise_seed=6100 + i)
        p = single if mode == "single" else grid
        spots = [p(s, flat).cpu().numpy() for s in chans]
        pl, mi = dapi(da)
        want.append((spots, pl.cpu().numpy(), mi.cpu().numpy()))
        reqs.append(FovRequest(i, [s.cpu().pin_memory() for s in chans], flat.cpu().pin_memory(),
                               arrival=float(arrivals[i]), dapi=da.cpu().pin_memory(), psf=mode, fov=fov))
        del chans, da
    srv = HybServer({"single": single, "grid": grid}, batch_tiles={"single": 36, "grid": 50}, slots=2, dapi=dapi)
    done = srv.run(reqs)
    st = srv.stats
    log(f"served stream: {st['fovs_per_s']:.2f} FOV/s, p50 {st['p50_latency_s']:.3f} s, p99 {st['p99_latency_s']:.3f} s, "
        f"tiles {st['tiles_per_mode']}, drift blocks {st['drift_blocks']}")
    assert st["failed"] == 0 and st["drift_blocks"] == 3 and st["dapi_stacks"] == 6
    for r in done:
        spots, pl, mi = want[r.rid]
        assert all(np.array_equal(a, b) for a, b in zip(r.spots, spots)) and len(spots[0]) > 1000, r.rid
        assert np.array_equal(r.Xh_plus, pl) and np.array_equal(r.Xh_minus, mi), r.rid

    def feats(i):
        return types.SimpleNamespace(Xh_plus=torch.from_numpy(want[i][1]).to(dev),
                                     Xh_minus=torch.from_numpy(want[i][2]).to(dev))
    for members in ([0, 1], [2, 3], [4, 5]):
        ref = members[len(members) // 2]
        dual = DualAligner(feats(ref), th=4)
        for i in members:
            f, tp, tm, n_p, n_m = dual.get_best_translation_pointsV2(feats(i))
            g = done[i].drift
            assert done[i].drift_ref == ref and np.array_equal(f, g[0]) and (n_p, n_m) == (g[3], g[4]), i
        # the reference member is the shifted one: the other member's drift is the negated shift's opposite
        sh = np.array(plan[ref][3]) - np.array(plan[members[0]][3])
        f0 = done[members[0]].drift[0]
        log(f"block {plan[ref][0]}: drift of member {members[0]} {f0}, shift {sh}")
        assert np.array_equal(np.round(f0).astype(int), sh), (f0, sh)
