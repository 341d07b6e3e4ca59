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
