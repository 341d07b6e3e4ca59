"""CPU tests: the oracle against the reference's known answers, golden fixtures,
brute force and closed forms (no GPU)."""
import numpy as np
import pytest

from oracle import align as oalign
from oracle import blur as oblur
from oracle import deconv as odec
from oracle import fill as ofill
from oracle import maxima as omax


# ------------------------------------------------------------------ blur
def test_blur_reference_known_answers():
    """tests/test_blur.py:62-100 of the reference."""
    arr = np.zeros((3, 3, 3), np.float32)
    arr[0, 0, 0] = 1
    third = np.float32(0.33333334)
    for axis, pos in ((0, (1, 0, 0)), (1, (0, 1, 0)), (2, (0, 0, 1))):
        want = np.zeros((3, 3, 3), np.float32)
        want[0, 0, 0] = 0.5
        want[pos] = third
        assert np.array_equal(oblur.box_1d(arr, 2, axis=axis), want)


@pytest.mark.parametrize("shape,axis,size", [((5, 5), 0, 3), ((5, 5), 1, 3), ((4, 4, 4), 0, 3), ((4, 4, 4), 1, 5),
                                             ((4, 4, 4), 2, 3), ((3, 6, 7), 1, 31), ((2, 5, 9), -1, 4)])
def test_blur_vs_bruteforce_and_numpy_reference(shape, axis, size):
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(shape) * 100).astype(np.float32)
    got = oblur.box_1d(a, size, axis=axis)
    ax = a.ndim - 1 if axis < 0 else axis
    assert np.array_equal(got, oblur.box_1d_bruteforce(a, size, axis=ax))
    # the reference test's numpy_box_1d (clipped-window mean, tests/test_blur.py:9-23)
    d = size // 2
    want = np.empty_like(a)
    for idx in np.ndindex(a.shape):
        sl = list(idx)
        sl[ax] = slice(max(0, idx[ax] - d), min(a.shape[ax], idx[ax] + d + 1))
        want[idx] = a[tuple(sl)].mean()
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4)


def test_box_errors():
    a = np.ones((3, 3, 3), np.float32)
    with pytest.raises(ValueError):
        oblur.box(a, 3, axes=(0, 1, 2, 3))
    with pytest.raises(TypeError):
        oblur.box(a, 3, axes="bad")
    with pytest.raises(ValueError):
        oblur.box(a, (3, 3), axes=(0, 1, 2))


# ------------------------------------------------------------------ fill
def test_reflect_reference_known_answers():
    """tests/test_fill.py:22-133 of the reference."""
    a = np.array([1, 2, 3, 4, 5], np.float32)
    assert ofill.reflect(a, 2, mode="out").tolist() == [1, 2, 3, 2, 1]
    assert ofill.reflect(a, 2, mode="in").tolist() == [5, 4, 3, 4, 5]
    d2 = np.arange(1, 16, dtype=np.float32).reshape(3, 5)
    assert ofill.reflect(d2, 1, axis=0, mode="in").tolist() == [[11, 12, 13, 14, 15], [6, 7, 8, 9, 10],
                                                                [11, 12, 13, 14, 15]]
    assert ofill.reflect(d2, 3, axis=1, mode="in").tolist() == [[1, 2, 5, 4, 5], [6, 7, 10, 9, 10],
                                                                [11, 12, 15, 14, 15]]
    assert ofill.reflect(d2, 1, axis=0, mode="out").tolist() == [[1, 2, 3, 4, 5], [6, 7, 8, 9, 10], [1, 2, 3, 4, 5]]
    with pytest.raises(IndexError):
        ofill.reflect(a, 5)
    with pytest.raises(ValueError):
        ofill.reflect(a, 1, mode="bad")
    with pytest.raises(ValueError, match="output array shape"):
        ofill.reflect(a, 1, out=np.zeros(4, np.float32))


# ------------------------------------------------------------------ maxima
@pytest.mark.parametrize("shape,delta", [((6, 7, 8), 1), ((5, 6, 7), 2), ((7, 7, 7), 3), ((2, 2, 2), 1),
                                         ((10, 5, 3), 1)])
def test_maxima_vs_bruteforce(shape, delta):
    rng = np.random.default_rng(1)
    img = rng.random(shape).astype(np.float32)
    z, x, y = omax.local_maxima(img, np.float32(0.5), delta)
    assert list(zip(z.tolist(), x.tolist(), y.tolist())) == omax.local_maxima_bruteforce(img, 0.5, delta)


def test_maxima_reference_behaviour():
    """tests/test_maxima.py:26-129 of the reference."""
    img = np.zeros((5, 5, 5), np.float32)
    img[1:4, 1:4, 1:4] = 0.5
    img[2, 2, 2] = 1.0
    r = omax.find_local_maxima(img, 0.8, 1, 1, raw=img)
    assert r.shape == (1, 8) and np.allclose(r[0, :3], 2.0)
    img = np.zeros((5, 5, 5), np.float32)
    img[1, 1, 1], img[3, 3, 3] = 0.9, 0.6
    assert omax.find_local_maxima(img, 0.5, 1, 1, raw=img).shape[0] == 2
    assert omax.find_local_maxima(img, 0.8, 1, 1, raw=img).shape[0] == 1
    assert omax.find_local_maxima(img, 1.5, 1, 1, raw=img).shape == (0, 8)
    with pytest.raises(TypeError):
        omax.find_local_maxima(img, 0.5, 6, 1, raw=img)
    # a NaN threshold (threshold=None in the reference) rejects nothing
    img2 = np.zeros((3, 3, 3), np.float32)
    img2[1, 1, 1] = 1
    z, x, y = omax.local_maxima(img2, np.float32(np.nan), 1)
    assert list(zip(z.tolist(), x.tolist(), y.tolist())) == omax.local_maxima_bruteforce(img2, np.nan, 1)
    assert len(z) == 27 - 6     # the peak plus every zero voxel not touching it (ties are maxima)


def test_fit_vectorised_vs_scalar_and_closed_forms():
    rng = np.random.default_rng(2)
    img = (rng.random((9, 12, 13)) * 100).astype(np.float32)
    raw = (rng.random((9, 12, 13)) * 1000).astype(np.float32)
    pts = [(0, 0, 0), (4, 6, 6), (8, 11, 12), (1, 5, 0), (3, 0, 7)]
    for dfit in (1, 3, 5, 6):
        v = omax.delta_fit(img, raw, *map(np.array, zip(*pts)), dfit, 1.0, 1.5)
        for k, p in enumerate(pts):
            np.testing.assert_allclose(v[k], omax.delta_fit_scalar(img, raw, *p, dfit, 1.0, 1.5), rtol=1e-10,
                                       atol=1e-9)
    # symmetric peak -> integer centroid; background = min sample
    sym = np.zeros((9, 9, 9), np.float32)
    zz, xx, yy = np.meshgrid(*[np.arange(9) - 4] * 3, indexing="ij")
    sym[:] = np.exp(-(zz ** 2 + xx ** 2 + yy ** 2) / 4.0)
    r = omax.delta_fit(sym, sym, np.array([4]), np.array([4]), np.array([4]), 3, 1.0, 1.5)[0]
    assert np.allclose(r[:3], 4.0) and r[3] == sym[np.abs(zz) ** 2 + xx ** 2 + yy ** 2 <= 9].min()
    # constant image -> hn = a = 0 (std replaced by 1), centroid falls back to the candidate
    c = np.full((5, 5, 5), 7.0, np.float32)
    r = omax.delta_fit(c, c, np.array([2]), np.array([2]), np.array([2]), 2, 1.0, 1.5)[0]
    assert r[4] == 0 and r[6] == 0 and np.allclose(r[:3], 2)
    # one sample: std_G = 0 -> NaN, as in the reference kernel
    r = omax.delta_fit(c, c, np.array([2]), np.array([2]), np.array([2]), 0, 1.0, 1.5)[0]
    assert np.isnan(r[4]) and np.isnan(r[6])


# ------------------------------------------------------------------ deconvolution
def test_padding_maps_match_sequential_reflections():
    for sz, zpad in [(35, 35), (40, 39), (40, 19), (20, 19), (5, 1), (3, 7), (1, 3)]:
        Zp = sz + 2 * zpad
        a = np.zeros(Zp)
        a[zpad:zpad + sz] = np.arange(1, sz + 1)
        ofill.reflect(a, zpad, mode="in", axis=0, out=a)
        ofill.reflect(a, zpad + sz - 1, mode="out", axis=0, out=a)
        m = odec.z_map(sz, zpad)
        assert np.array_equal(np.where(m >= 0, m + 1, 0), a), (sz, zpad)
    for T, ov in [(678, 89), (478, 89), (40, 5)]:
        ts = T - 2 * ov
        for d in range(1, T + 1):
            for first in (True, False):
                start = ov if first else 0
                if (first and d > ts + ov) or start + d > T:
                    continue
                a = np.zeros(T)
                a[start:start + d] = np.arange(1, d + 1)
                if first:
                    ofill.reflect(a, ov, mode="in", axis=0, out=a)
                elif d < T:
                    ofill.reflect(a, d - 1, mode="out", axis=0, out=a)
                m = odec.padding_maps(T, start, d, first)
                assert np.array_equal(np.where(m >= 0, m + 1, 0), a)


def test_laplacian_closed_form_and_wiener_delta():
    for shp in [(5, 5, 5), (14, 40, 40), (7, 9, 11)]:
        L = np.fft.fftn(odec.laplacian_3d(shp).astype(np.float64))
        assert np.allclose(np.abs(L) ** 2, odec.laplacian_power(shp))
        assert abs(L[0, 0, 0] - 1) < 1e-12            # stencil sums to 1
    sz, ts, ov, zpad, beta = 4, 10, 3, 2, 1e-2
    psf = np.zeros((sz, ts, ts), np.float32)
    psf[sz // 2, ts // 2, ts // 2] = 1
    W = odec.wiener(psf, sz, ts, zpad, ov, beta)[0]
    assert np.allclose(W, 1 / (1 + beta * odec.laplacian_power(W.shape)))


def test_blur_fold_identity():
    """Deconvolving with W * (1 - b(kx) b(ky)), b the (2D+1)-tap mean's transfer function on the T-periodic tile
    (summed term by term), gives the reference's deconvolution minus its clipped box blur (deconvolver.py:160-163)
    wherever the boxes stay inside the tile, to float64 rounding; near the tile edges, where the reference clips the
    boxes, the two differ.  The closed form the HIP build uses (Dirichlet kernel) equals the term-by-term sum."""
    from mermake_amd.synth import gaussian_psf, stack_np
    sz, ts, ov, zpad, D = 8, 40, 14, 7, 5
    raw, _ = stack_np(5, (sz, 68, 68), 30)
    psf = gaussian_psf((sz, 15, 15), 1.5, 1.2)
    W = odec.wiener(psf, sz, ts, zpad, ov, 1e-3)[0]
    T = W.shape[1]
    k = np.arange(T)
    m = np.arange(-D, D + 1)
    b = np.cos(2 * np.pi * np.outer(k, m) / T).sum(1) / (2 * D + 1)
    with np.errstate(invalid="ignore", divide="ignore"):
        closed = np.sin(np.pi * (2 * D + 1) * k / T) / ((2 * D + 1) * np.sin(np.pi * k / T))
    closed[0] = 1.0
    assert np.abs(closed - b).max() < 1e-12
    Wf = W * (1.0 - b[None, :, None] * b[None, None, :])
    blur = list(odec.tile_wise(raw, W, None, 2 * D, ts, ov, zpad))
    fold = list(odec.tile_wise(raw, Wf, None, None, ts, ov, zpad))
    for (x, y, a, _), (_, _, f, _) in zip(blur, fold):
        win = (slice(None), slice(D, min(a.shape[1], T - D)), slice(D, min(a.shape[2], T - D)))
        scale = np.abs(a[win]).max()
        assert np.abs(a[win] - f[win]).max() <= 1e-12 * scale, (x, y)
        assert np.abs(a - f).max() > 1e-6 * scale, (x, y)   # the clipped edge boxes are not periodic


def test_center_psf_and_tiled():
    """tests/test_deconvolver.py:158-218 of the reference."""
    for shp in [(20, 50, 50), (10, 25, 25), (30, 75, 75)]:
        r = odec.center_psf(np.ones(shp), 20, 50)
        assert r.shape == (20, 50, 50) and np.isclose(r.sum(), 1.0)
    assert odec.center_psf(np.ones((10, 25, 25)), 20, 50)[10, 25, 25] > 0
    pos = [(x, y) for x, y, *_ in odec.tiled((20, 100, 100), 50, 10)]
    assert pos == [(0, 0), (0, 50), (50, 0), (50, 50)]


def test_oracle_reproduces_golden_fixture_tiles():
    """The reference's float16 golden of full_deconv on the real fixture (tests/test_deconvolver.py:422-434)."""
    from fixtures import OV, TS, fixture_stack, golden_accept, psf_750
    stack, wins, gz = fixture_stack()
    zpad = stack.shape[0]
    W = odec.wiener(psf_750(), stack.shape[0], TS, zpad, OV, 0.001, mode="fp64")
    seen = 0
    for x, y, t, _ in odec.tile_wise(stack, W, None, None, TS, OV, zpad, only=set(wins)):
        own = t[:, OV:OV + TS, OV:OV + TS][list(gz)]
        exact, worst = golden_accept(own, wins[(x, y)][2])
        assert exact > 0.995 and worst <= 1.0, (x, y, exact, worst)
        seen += 1
    assert seen == 3


# ------------------------------------------------------------------ drift
def test_drift_known_answer():
    """drift pickle of the reference fixture: N+ = 621, N- = 261, zero drift."""
    import types
    from fixtures import dapi_features, drift_known_answer
    xp, xm = dapi_features()
    ka = drift_known_answer()
    ref = types.SimpleNamespace(Xh_plus=xp, Xh_minus=xm)
    f, tp, tm, n_p, n_m = oalign.DualAligner(ref, th=4).get_best_translation_pointsV2(ref)
    assert (n_p, n_m) == (ka["N_plus"], ka["N_minus"]) and np.allclose(f, 0)
    # 621 = rows with h > 4 and y > 4 (the reference set is thresholded on column 2)
    assert int(((xp[:, 7] > 4) & (xp[:, 2] > 4)).sum()) == 621


def test_drift_combination_rules():
    """tests/test_align.py:181-233 of the reference."""
    assert np.allclose(oalign.combine([1, 2, 3], 0, [1, 2, 3], 0)[0], 0)
    assert np.allclose(oalign.combine([1, 2, 3], 10, [1, 2, 3], 5)[0], [-1, -2, -3])
    assert np.allclose(oalign.combine([1, 2, 3], 10, [10, 20, 30], 5)[0], [-1, -2, -3])
    assert np.allclose(oalign.combine([1, 2, 3], 3, [10, 20, 30], 5)[0], [-10, -20, -30])


def test_drift_integer_shift_oracle():
    rng = np.random.default_rng(3)
    ref = rng.uniform((0, 60, 60), (40, 1400, 1400), size=(600, 3)).astype(np.float32)
    for shift in [(2, -15, 25), (-3, 30, -30)]:
        mov = ref + np.array(shift, np.float32)
        t = oalign.Aligner(ref).get_best_translation_points(mov)
        assert np.array_equal(np.round(t).astype(int), shift)


def test_empty_reference_raises():
    with pytest.raises(ValueError):
        oalign.Aligner(np.ones((0, 4)), resc=1, trim=1)


# ------------------------------------------------------------------ pins of the fit's correlations
def _sphere_mask(d):
    """Offsets (dz, dx, dy) of the delta_fit sphere from np.indices (order-free)."""
    g = np.indices((2 * d + 1,) * 3) - d
    m = (g ** 2).sum(0) <= d * d
    return g[0][m], g[1][m], g[2][m]


def _gauss_field(shape, c, sz, sxy):
    z, x, y = np.indices(shape)
    return np.exp(-((z - c[0]) ** 2 / (2 * sz * sz) + (x - c[1]) ** 2 / (2 * sxy * sxy)
                    + (y - c[2]) ** 2 / (2 * sxy * sxy)))


@pytest.mark.parametrize("dfit,sz,sxy", [(3, 1.0, 1.5), (5, 1.0, 1.5), (2, 1.3, 0.8)])
def test_fit_correlations_closed_forms(dfit, sz, sxy):
    """``a`` and ``hn`` (maxima.cu:249-298) are Pearson correlations with the sampled Gaussian, computed
    with population statistics.  Closed forms: an image equal to c*G + d on the sphere gives hn = 1
    (c > 0) or -1 (c < 0), a raw equal to c*G + d gives a = +-1 exactly; a ddof=1 std gives
    (n-1)/n instead, swapped columns flip the signs.  For arbitrary volumes the correlations equal
    numpy.corrcoef of the sphere samples (interior candidates: no reflection)."""
    shape = (2 * dfit + 5, 2 * dfit + 7, 2 * dfit + 6)
    c = (dfit + 2, dfit + 3, dfit + 2)
    G = _gauss_field(shape, c, sz, sxy)
    img = 3.0 * G + 100.0
    raw = -2.0 * G + 50.0
    r = omax.delta_fit(img, raw, np.array([c[0]]), np.array([c[1]]), np.array([c[2]]), dfit, sz, sxy)[0]
    assert abs(r[6] - 1.0) < 1e-12 and abs(r[4] + 1.0) < 1e-12, r
    r = omax.delta_fit(-img, -raw, np.array([c[0]]), np.array([c[1]]), np.array([c[2]]), dfit, sz, sxy)[0]
    assert abs(r[6] + 1.0) < 1e-12 and abs(r[4] - 1.0) < 1e-12, r
    # arbitrary volumes: numpy.corrcoef of the samples
    rng = np.random.default_rng(dfit)
    img = rng.random(shape) * 50 + 3 * G
    raw = rng.random(shape) * 900 - G * 200
    dz, dx, dy = _sphere_mask(dfit)
    gs = np.exp(-(dz ** 2 / (2 * sz * sz) + dx ** 2 / (2 * sxy * sxy) + dy ** 2 / (2 * sxy * sxy)))
    iv = img[c[0] + dz, c[1] + dx, c[2] + dy]
    rv = raw[c[0] + dz, c[1] + dx, c[2] + dy]
    r = omax.delta_fit(img, raw, np.array([c[0]]), np.array([c[1]]), np.array([c[2]]), dfit, sz, sxy)[0]
    np.testing.assert_allclose(r[6], np.corrcoef(iv, gs)[0, 1], rtol=1e-12)
    np.testing.assert_allclose(r[4], np.corrcoef(rv, gs)[0, 1], rtol=1e-12)
    assert r[3] == iv.min() and r[5] == raw[c] and r[7] == img[c]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fit_centroid_linear_ramp(axis):
    """On a linear ramp alpha*coord + beta the background is the sample at offset -d, so the weights are
    alpha*(off + d) and the centroid moves by mean(off^2)/d along the ramp's axis, 0 along the others."""
    d = 3
    shape = (13, 14, 15)
    c = (6, 7, 7)
    img = 2.5 * np.indices(shape)[axis] + 40.0
    r = omax.delta_fit(img, img, np.array([c[0]]), np.array([c[1]]), np.array([c[2]]), d, 1.0, 1.5)[0]
    off = _sphere_mask(d)[axis]
    want = np.array(c, dtype=np.float64)
    want[axis] += (off.astype(np.float64) ** 2).mean() / d
    np.testing.assert_allclose(r[:3], want, rtol=0, atol=1e-12)


# ------------------------------------------------------------------ fftconvolve vs SciPy
@pytest.mark.parametrize("sa,sb", [((17,), (5,)), ((9, 12), (4, 3)), ((5, 17, 23), (3, 9, 4)), ((6, 8, 7), (6, 8, 7)),
                                   ((11, 4, 9), (2, 4, 1))])
def test_fftconvolve_matches_scipy(sa, sb):
    """``utils.py:10-36`` in all three modes against ``scipy.signal.fftconvolve`` (rank 1 to 3; for 'valid'
    the first input is at least as large as the second in every axis, where SciPy does not swap them)."""
    from scipy.signal import fftconvolve as sfc
    rng = np.random.default_rng(len(sa) * 10 + sa[0])
    a = rng.standard_normal(sa)
    b = rng.standard_normal(sb)
    for mode in ("full", "same", "valid"):
        got = oalign.fftconvolve(a, b, mode=mode)
        want = sfc(a, b, mode=mode)
        assert got.shape == want.shape, (mode, got.shape, want.shape)
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    with pytest.raises(ValueError):
        oalign.fftconvolve(a, b, mode="nope")


# ------------------------------------------------------------------ flat field by linearity
def test_flat_field_divide_and_multiply_back_by_linearity():
    """deconvolver.py:127-129,166-168: the data mask of the padded tile is divided by the flat before the
    reflections, and only the padded tile (the raw view) is multiplied back after the blur.  So

    * with or without blur, deconvolving v with flat f equals deconvolving the float image v/f without a
      flat (the deconvolved tile stays flat-divided);
    * the raw views hold (v/f)*f on the data mask and the reflected v/f in the pads."""
    from mermake_amd.synth import gaussian_psf
    rng = np.random.default_rng(5)
    sz, sx, sy, ts, ov, zpad = 6, 70, 64, 40, 9, 5
    v = rng.poisson(300.0, size=(sz, sx, sy)).astype(np.uint16)
    f = (0.6 + 0.4 * rng.random((sx, sy))).astype(np.float64)
    psf = gaussian_psf((sz, 9, 9), 1.2, 1.1)
    W = odec.wiener(psf, sz, ts, zpad, ov, 1e-3, mode="fp64")
    vf = v.astype(np.float64) / f[None]
    for blur in (None, 6):
        a = list(odec.tile_wise(v, W, f, blur, ts, ov, zpad))
        b = list(odec.tile_wise(vf, W, None, blur, ts, ov, zpad))
        assert len(a) == len(b) == 4
        for (x, y, ra, pa), (_, _, rb, pb) in zip(a, b):
            xs, ys = (ov if x == 0 else 0), (ov if y == 0 else 0)
            x0, y0 = max(x - ov, 0), max(y - ov, 0)
            xd, yd = min(x + ts + ov, sx) - x0, min(y + ts + ov, sy) - y0
            m = np.zeros(ra.shape[1:], bool)
            m[xs:xs + xd, ys:ys + yd] = True
            fm = np.ones(ra.shape[1:])
            fm[xs:xs + xd, ys:ys + yd] = f[x0:x0 + xd, y0:y0 + yd]
            np.testing.assert_allclose(ra, rb, rtol=1e-9, atol=1e-9 * np.abs(rb).max())
            # raw views: (v/f)*f on the data mask, v/f (reflected) elsewhere
            np.testing.assert_allclose(pa[:, m], (pb * fm[None])[:, m], rtol=1e-14)
            np.testing.assert_allclose(pa[:, ~m], pb[:, ~m], rtol=1e-14)
            assert np.allclose(pa[:, m], v[:, x0:x0 + xd, y0:y0 + yd].reshape(sz, -1), rtol=1e-13)
