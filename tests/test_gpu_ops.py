"""GPU parity tests of the operators (HIP path vs the oracle), through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import blur as oblur
from oracle import deconv as odec
from oracle import fill as ofill
from oracle import maxima as omax

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


# ------------------------------------------------------------------ blur
def test_box_1d_reference_known_answers(dev):
    """tests/test_blur.py:62-100 of the reference: exact 3x3x3 values for size 2."""
    from mermake_amd import blur
    arr = torch.zeros((3, 3, 3), dtype=torch.float32, device=dev)
    arr[0, 0, 0] = 1
    third = np.float32(1) / np.float32(3)
    want0 = np.zeros((3, 3, 3), np.float32); want0[0, 0, 0] = 0.5; want0[1, 0, 0] = third
    want1 = np.zeros((3, 3, 3), np.float32); want1[0, 0, 0] = 0.5; want1[0, 1, 0] = third
    want2 = np.zeros((3, 3, 3), np.float32); want2[0, 0, 0] = 0.5; want2[0, 0, 1] = third
    for axis, want in ((0, want0), (1, want1), (2, want2)):
        got = blur.box_1d(arr, 2, axis=axis).cpu().numpy()
        assert np.array_equal(got, want), axis
    with pytest.raises(ValueError):
        blur.box_1d(arr, 2, axis=0, out=arr)


@pytest.mark.parametrize("shape,axis,size", [((5, 5), 0, 3), ((5, 5), 1, 3), ((4, 4, 4), 0, 3), ((4, 4, 4), 1, 5),
                                             ((4, 4, 4), 2, 3), ((7, 33, 40), 1, 31), ((7, 33, 40), 2, 51),
                                             ((3, 50, 61), 1, 30), ((6, 9, 8), -1, 4)])
def test_box_1d_bitexact(dev, shape, axis, size):
    from mermake_amd import blur
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(shape) * 1000).astype(np.float32)
    got = blur.box_1d(T(a, dev), size, axis=axis).cpu().numpy()
    want = oblur.box_1d(a, size, axis=axis)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("T_,delta,scale", [(678, 15, 1000.0), (100, 15, 1.0), (40, 15, 1e-3), (17, 15, 1000.0),
                                             (300, 15, 10.0), (200, 15, 1.0), (1000, 15, 100.0),
                                             (678, 25, 1000.0), (478, 25, 1.0), (53, 25, 1.0), (31, 25, 1e-30)])
def test_blur_sub_fused_bitexact(dev, T_, delta, scale):
    """The fused blur-subtract (deconvolver.py:160-163) equals res - box_1d(box_1d(res, axis 1), axis 2) bit for bit,
    including windows clipped at the borders, planes narrower than the window and tiny sums."""
    from mermake_amd import _lib as L
    rng = np.random.default_rng(T_ + delta)
    nplanes = 3
    res = (rng.standard_normal((nplanes, T_, T_)) * scale).astype(np.float32)
    res[0, :5, :5] = 0.0                     # zero windows
    res[1, 3, :] = -0.0
    size = 2 * delta
    want = res - oblur.box_1d(oblur.box_1d(res, size, axis=1), size, axis=2)
    r = T(res, dev)
    out = torch.empty_like(r)
    L.call("mm_blur_sub_xy", L.ptr(r), L.ptr(out), None, nplanes, T_, delta, L.stream_ptr(dev))
    got = out.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("T_,delta,lo,hi", [(678, 15, 82, 596), (678, 15, 1, 677), (100, 15, 23, 77), (100, 15, 40, 41),
                                           (200, 25, 10, 190), (60, 15, 5, 55)])
def test_blur_sub_window_bitexact(dev, T_, delta, lo, hi):
    """The windowed blur-subtract equals the whole-plane one at rows and columns [lo, hi) and leaves the rest of
    the output untouched."""
    from mermake_amd import _lib as L
    rng = np.random.default_rng(T_ + lo)
    nplanes = 2
    res = (rng.standard_normal((nplanes, T_, T_)) * 100).astype(np.float32)
    size = 2 * delta
    want = res - oblur.box_1d(oblur.box_1d(res, size, axis=1), size, axis=2)
    r = T(res, dev)
    out = torch.full_like(r, 12345.0)
    L.call("mm_blur_sub_xy_window", L.ptr(r), L.ptr(out), None, nplanes, T_, delta, lo, hi, L.stream_ptr(dev))
    got = out.cpu().numpy()
    assert np.array_equal(got[:, lo:hi, lo:hi].view(np.uint32), want[:, lo:hi, lo:hi].view(np.uint32))
    inside = np.zeros((T_, T_), bool)
    inside[lo:hi, lo:hi] = True
    assert np.all(got[:, ~inside] == 12345.0)
    with pytest.raises(L.MermakeLibraryError):
        L.call("mm_blur_sub_xy_window", L.ptr(r), L.ptr(out), None, nplanes, T_, delta, 5, T_ + 1, L.stream_ptr(dev))


def test_box_multi_axis(dev):
    from mermake_amd import blur
    a = np.arange(64, dtype=np.float32).reshape(4, 4, 4)
    for axes, size in (((0, 1), 3), ((1, 2), (3, 5))):
        got = blur.box(T(a, dev), size, axes=axes).cpu().numpy()
        assert np.array_equal(got, oblur.box(a, size, axes=axes))
        got2 = blur.box_2d(T(a, dev), size, axes=axes).cpu().numpy()
        assert np.array_equal(got2, got)
    with pytest.raises(ValueError):
        blur.box(T(a, dev), 3, axes=(0, 1, 2, 3))
    with pytest.raises(TypeError):
        blur.box(T(a, dev), 3, axes="bad")
    with pytest.raises(ValueError):
        blur.box(T(a, dev), (3, 3), axes=(0, 1, 2))


# ------------------------------------------------------------------ fill
def test_reflect_reference_known_answers(dev):
    """tests/test_fill.py:22-133 of the reference."""
    from mermake_amd.fill import reflect
    a1 = torch.tensor([1, 2, 3, 4, 5], dtype=torch.float32, device=dev)
    assert reflect(a1, 2, mode="out").tolist() == [1, 2, 3, 2, 1]
    assert reflect(a1, 2, mode="in").tolist() == [5, 4, 3, 4, 5]
    assert a1.tolist() == [1, 2, 3, 4, 5]
    d2 = torch.arange(1, 16, dtype=torch.float32, device=dev).reshape(3, 5)
    assert reflect(d2, 2, axis=0, mode="in").tolist() == d2.tolist()
    assert reflect(d2, 1, axis=0, mode="in").tolist() == [[11, 12, 13, 14, 15], [6, 7, 8, 9, 10], [11, 12, 13, 14, 15]]
    assert reflect(d2, 2, axis=1, mode="in").tolist() == [[5, 4, 3, 4, 5], [10, 9, 8, 9, 10], [15, 14, 13, 14, 15]]
    assert reflect(d2, 3, axis=1, mode="in").tolist() == [[1, 2, 5, 4, 5], [6, 7, 10, 9, 10], [11, 12, 15, 14, 15]]
    assert reflect(d2, 0, axis=0, mode="out").tolist() == d2.tolist()
    assert reflect(d2, 1, axis=0, mode="out").tolist() == [[1, 2, 3, 4, 5], [6, 7, 8, 9, 10], [1, 2, 3, 4, 5]]
    assert reflect(d2, 2, axis=1, mode="out").tolist() == [[1, 2, 3, 2, 1], [6, 7, 8, 7, 6], [11, 12, 13, 12, 11]]
    x = a1.clone()
    assert reflect(x, 2, mode="out", out=x) is x and x.tolist() == [1, 2, 3, 2, 1]
    with pytest.raises(ValueError, match="output array shape"):
        reflect(a1, 1, out=torch.zeros(4, device=dev))
    with pytest.raises(ValueError, match="output array dtype"):
        reflect(a1, 1, out=torch.zeros(5, dtype=torch.int32, device=dev))
    with pytest.raises(IndexError):
        reflect(a1, 5)
    with pytest.raises(ValueError):
        reflect(a1, 1, mode="invalid")
    with pytest.raises(IndexError):
        reflect(a1, 1, axis=1)


def test_reflect_random_vs_oracle(dev):
    from mermake_amd.fill import reflect
    rng = np.random.default_rng(2)
    a = rng.random((4, 7, 9)).astype(np.float32)
    for axis in range(3):
        for i in range(a.shape[axis]):
            for mode in ("in", "out"):
                got = reflect(T(a, dev), i, axis=axis, mode=mode).cpu().numpy()
                assert np.array_equal(got, ofill.reflect(a, i, axis=axis, mode=mode)), (axis, i, mode)
    # strided (non-contiguous) input
    t = T(a, dev).transpose(0, 2)
    got = reflect(t, 2, axis=1, mode="in").cpu().numpy()
    assert np.array_equal(got, ofill.reflect(a.transpose(2, 1, 0), 2, axis=1, mode="in"))


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("shape", [(1, 1, 8), (3, 5, 7), (2, 6, 16), (4, 29, 59), (2, 113, 3), (1, 13, 118),
                                   (2, 3, 678), (2, 478, 2), (3, 11, 105), (1, 2, 1101)])
def test_fft3d_vs_numpy(dev, shape):
    from mermake_amd.deconvolver import fftn_c2c
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    got = fftn_c2c(T(x, dev)).cpu().numpy()
    want = np.fft.fftn(x.astype(np.complex128))
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 2e-6
    inv = fftn_c2c(T(x, dev), inverse=True).cpu().numpy()
    want_i = np.fft.ifftn(x.astype(np.complex128))
    assert np.linalg.norm(inv - want_i) / np.linalg.norm(want_i) < 2e-6


@pytest.mark.parametrize("shape,box", [((18, 40, 36), (9, 20)), ((5, 7, 9), (5, 7)), ((4, 30, 26), (1, 1)),
                                       ((6, 24, 60), (3, 11)), ((3, 5, 8), (0, 4))])
def test_fft3d_box_equals_full_transform(dev, shape, box):
    """mm_fft3d_c2c_box on a grid that is zero outside [0, m0) x [0, m1) gives mm_fft3d_c2c's forward transform of
    the whole grid bit for bit (the skipped lines are zero before their pass), and numpy's within float32."""
    from mermake_amd import _lib as L
    rng = np.random.default_rng(sum(shape))
    x = np.zeros(shape, np.complex64)
    m0, m1 = box
    x[:m0, :m1, :] = (rng.standard_normal((m0, m1, shape[2])) + 1j * rng.standard_normal((m0, m1, shape[2])))
    a, b = T(x, dev), T(x, dev)
    st = L.stream_ptr(dev)
    L.call("mm_fft3d_c2c", L.ptr(a), *shape, 7, 0, st)
    L.call("mm_fft3d_c2c_box", L.ptr(b), *shape, m0, m1, st)
    assert torch.equal(a, b)
    if m0 and m1:
        want = np.fft.fftn(x.astype(np.complex128))
        assert np.linalg.norm(b.cpu().numpy() - want) / np.linalg.norm(want) < 2e-6


# ------------------------------------------------------------------ maxima + fit
def _peaky(rng, shape, n=30):
    img = rng.normal(0, 1, shape).astype(np.float32)
    for _ in range(n):
        z, x, y = [rng.integers(0, s) for s in shape]
        img[z, x, y] += rng.uniform(5, 20)
    return img


@pytest.mark.parametrize("shape,delta", [((12, 20, 24), 1), ((10, 17, 13), 2), ((8, 30, 31), 3), ((12, 12, 12), 5),
                                         ((2, 2, 2), 1), ((1, 1, 1), 1), ((10, 5, 3), 1), ((16, 40, 70), 4),
                                         ((14, 33, 90), 5), ((20, 64, 130), 3)])
def test_maxima_sets_exact(dev, shape, delta):
    from mermake_amd.maxima import maxima_coords
    rng = np.random.default_rng(4)
    img = _peaky(rng, shape)
    thr = np.float32(2.0)
    got, _ = maxima_coords(T(img, dev), float(thr), delta)
    got = got.cpu().numpy()
    z, x, y = omax.local_maxima(img, thr, delta)
    assert np.array_equal(got, np.stack([z, x, y], 1).astype(np.int32))
    if np.prod(shape) <= 4000:
        assert [tuple(r) for r in got] == omax.local_maxima_bruteforce(img, thr, delta)


@pytest.mark.parametrize("delta,sign", [(2, 1.0), (3, 1.0), (5, 1.0), (5, -1.0)])
def test_maxima_plateaus_exact(dev, delta, sign):
    """Values on a few levels: plateaus of equal neighbours (ties count as maxima) inside the
    volume and on its borders, for the distance-ordered sphere test and its reflecting form."""
    from mermake_amd.maxima import maxima_coords
    rng = np.random.default_rng(delta)
    img = rng.integers(0, 4, (14, 37, 75)).astype(np.float32)
    thr = 0.5 if sign > 0 else -2.5
    got, _ = maxima_coords(T(img, dev), thr, delta, sign=sign)
    z, x, y = omax.local_maxima(sign * img, np.float32(thr), delta)
    assert len(z) > 0
    assert np.array_equal(got.cpu().numpy(), np.stack([z, x, y], 1).astype(np.int32))


def test_maxima_ties_and_views(dev):
    from mermake_amd.maxima import find_local_maxima
    img = np.zeros((5, 6, 7), np.float32)
    img[2, 2, 2] = 1.0
    img[2, 2, 3] = 1.0      # tie: both are maxima (v >= neighbours)
    out = find_local_maxima(T(img, dev), threshold=0.5, delta=1, delta_fit=1, raw=T(img, dev))
    assert out.shape == (2, 8)
    big = np.zeros((5, 10, 12), np.float32)
    big[:, 1:7, 2:9] = img[:, :6, :7]
    view = T(big, dev)[:, 1:7, 2:9]    # strided view: reflection at the view border
    out2 = find_local_maxima(view, threshold=0.5, delta=1, delta_fit=1, raw=view)
    assert torch.equal(out.cpu(), out2.cpu())


def test_find_local_maxima_reference_behaviour(dev):
    """tests/test_maxima.py:26-166 of the reference."""
    from mermake_amd.maxima import find_local_maxima
    image = torch.zeros((5, 5, 5), dtype=torch.float32, device=dev)
    image[1:4, 1:4, 1:4] = 0.5
    image[2, 2, 2] = 1.0
    r = find_local_maxima(image, threshold=0.8, delta=1, delta_fit=1, raw=image.clone())
    assert r.shape == (1, 8) and all(abs(float(r[0, k]) - 2.0) < 0.5 for k in range(3))
    image = torch.zeros((5, 5, 5), dtype=torch.float32, device=dev)
    image[1, 1, 1] = 0.9
    image[3, 3, 3] = 0.6
    assert find_local_maxima(image, threshold=0.5, delta=1, delta_fit=1, raw=image).shape[0] == 2
    assert find_local_maxima(image, threshold=0.8, delta=1, delta_fit=1, raw=image).shape[0] == 1
    r = find_local_maxima(image, threshold=1.5, delta=1, delta_fit=1, raw=image)
    assert r.shape == (0, 8)
    with pytest.raises(TypeError, match="Delta must be an less than or equal to 5"):
        find_local_maxima(image, threshold=0.5, delta=6, delta_fit=1, raw=image)
    ones = torch.ones((3, 3, 3), dtype=torch.float32, device=dev)
    r1 = find_local_maxima(ones, threshold=0.5, delta=1, delta_fit=1, raw=(ones * 65535).to(torch.uint16))
    r2 = find_local_maxima(ones, threshold=0.5, delta=1, delta_fit=1, raw=ones.double())
    assert r1.shape == r2.shape == (27, 8)


@pytest.mark.parametrize("dfit,sz,sxy", [(1, 1.0, 1.5), (3, 1.0, 1.5), (5, 1.0, 1.5), (2, 2.0, 2.0), (6, 1.0, 1.0)])
def test_fit_vs_oracle(dev, dfit, sz, sxy):
    from mermake_amd.maxima import find_local_maxima
    rng = np.random.default_rng(5)
    img = _peaky(rng, (14, 40, 45), 40) * 100 + 500
    raw = (img * 1.3 + rng.normal(0, 5, img.shape)).astype(np.float32)
    got = find_local_maxima(T(img, dev), threshold=900, delta=1, delta_fit=dfit, raw=T(raw, dev), sigmaZ=sz,
                            sigmaXY=sxy).cpu().numpy()
    want = omax.find_local_maxima(img, threshold=900, delta=1, delta_fit=dfit, raw=raw, sigmaZ=sz, sigmaXY=sxy)
    assert got.shape == want.shape and got.shape[0] > 10
    assert np.abs(got[:, :3] - want[:, :3]).max() < 1e-3                      # coordinates (px)
    assert np.array_equal(got[:, 3], want[:, 3].astype(np.float32))          # background = min sample
    assert np.array_equal(got[:, 5], want[:, 5].astype(np.float32))          # habs
    assert np.array_equal(got[:, 7], want[:, 7].astype(np.float32))          # h
    assert np.abs(got[:, [4, 6]] - want[:, [4, 6]]).max() < 1e-5             # correlations in [-1, 1]


def test_fit_nan_when_single_sample(dev):
    """delta_fit = 0: one sample, std_G = 0 -> a, hn NaN (maxima.cu:249-254)."""
    from mermake_amd.maxima import find_local_maxima
    img = np.zeros((5, 5, 5), np.float32)
    img[2, 2, 2] = 3.0
    r = find_local_maxima(T(img, dev), threshold=1.0, delta=1, delta_fit=0, raw=T(img, dev)).cpu().numpy()
    assert r.shape == (1, 8) and np.isnan(r[0, 4]) and np.isnan(r[0, 6])
    assert r[0, 0] == 2 and r[0, 1] == 2 and r[0, 2] == 2 and r[0, 7] == 3.0


# ------------------------------------------------------------------ deconvolution
def _rel_l2(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape,ts,ov,zpad,beta,flat,blur", [
    ((10, 50, 50), 30, 5, 2, 1e-3, False, None),
    ((10, 50, 50), 30, 5, 2, 1e-3, True, None),
    ((6, 37, 53), 20, 4, 3, 1e-2, True, 5),
    ((5, 20, 20), 50, 5, 2, 1e-3, False, None),     # single tile larger than the image
    ((4, 45, 45), 15, 2, 1, 1e-4, False, 3),
    ((3, 40, 40), 20, 7, 5, 1e-3, True, None),      # zpad >= sz: zero planes in the z padding
    ((4, 70, 70), 50, 8, 2, 1e-3, True, 5),         # T = 66 = 11 * 6: prime-first passes, radix-6 final stage
    ((4, 50, 50), 23, 5, 2, 1e-3, False, None),     # T = 33 = 11 * 3
    ((4, 60, 60), 30, 7, 2, 1e-3, True, None),      # T = 44 = 11 * 4
    ((3, 600, 600), 500, 89, 3, 1e-4, True, 15),    # T = 678 = 113 * 6, the full-FOV tile (Rader prime stage)
    ((3, 300, 300), 200, 13, 2, 1e-3, True, None),  # T = 226 = 113 * 2 (Rader, radix-2 final stage)
    ((3, 340, 340), 301, 19, 2, 1e-3, False, 5),    # T = 339 = 113 * 3
    ((2, 500, 500), 400, 26, 2, 1e-3, True, None),  # T = 452 = 113 * 4
    ((40, 120, 120), 80, 10, 39, 1e-3, True, None),  # 40 planes, zpad = sz - 1: the length-80 z pass
    ((40, 260, 260), 200, 13, 39, 1e-3, False, 15),  # ... with T = 226 (Rader, blocked spectra)
    ((17, 100, 100), 60, 8, 16, 1e-3, True, None),   # length-80 z pass with a run-time depth
    ((3, 400, 400), 300, 89, 2, 1e-3, True, 15),    # T = 478 = 239 * 2 (Rader 239, FFT-238 as 17 x 14)
    ((10, 330, 330), 300, 89, 9, 1e-3, False, 15),  # ... with the length-80 z pass
])
def test_deconv_vs_oracle(dev, shape, ts, ov, zpad, beta, flat, blur):
    from mermake_amd.deconvolver import Deconvolver
    rng = np.random.default_rng(6)
    img = rng.integers(100, 4000, size=shape).astype(np.uint16)
    psf = rng.random((5, 7, 7)).astype(np.float32) + np.float32(0.1)
    fl = (0.7 + 0.6 * rng.random(shape[1:])).astype(np.float32) if flat else None
    d = Deconvolver(psf, shape, tile_size=ts, zpad=zpad, overlap=ov, beta=beta)
    W = odec.wiener(psf, shape[0], ts, zpad, ov, beta, mode="fp64")
    got = list(d.tile_wise(T(img, dev), flat_field=None if fl is None else T(fl, dev), blur_radius=blur))
    want = list(odec.tile_wise(img, W, fl, blur, ts, ov, zpad, mode="fp64"))
    want32 = list(odec.tile_wise(img, odec.wiener(psf, shape[0], ts, zpad, ov, beta, mode="fp32"), fl, blur, ts, ov,
                                 zpad, mode="fp32"))
    assert len(got) == len(want)
    for (gx, gy, gt, gr), (wx, wy, wt, wr), (_, _, _, wr32) in zip(got, want, want32):
        assert (gx, gy) == (wx, wy) and tuple(gt.shape) == wt.shape and tuple(gr.shape) == wr.shape
        assert _rel_l2(gt.cpu().numpy(), wt) < 1e-5
        assert np.array_equal(gr.cpu().numpy(), wr32)          # raw views: float32 (v/f)*f semantics
    out = d.apply(T(img, dev), flat_field=None if fl is None else T(fl, dev), blur_radius=blur).cpu().numpy()
    ref = odec.apply(img, W, fl, blur, ts, ov, zpad, mode="fp64")
    written = np.zeros(shape, bool)
    for x, y, t, _ in odec.tile_wise(img, W, fl, blur, ts, ov, zpad, mode="fp64"):
        written[:, x:x + t.shape[1] - 2 * ov, y:y + t.shape[2] - 2 * ov] = True
    assert _rel_l2(out[written], ref[written]) < 1e-5


@pytest.mark.parametrize("shape,ts,ov,lo,hi", [((3, 300, 300), 200, 13, 5, 215),   # T = 226 (Rader, blocked spectra)
                                                 ((3, 760, 760), 500, 89, 67, 611),  # T = 678
                                                 ((3, 420, 420), 300, 89, 67, 411),  # T = 478 (Rader 239)
                                                 ((4, 120, 120), 60, 20, 9, 91)])    # T = 100
def test_deconv_window_bitexact(dev, shape, ts, ov, lo, hi):
    """mm_deconv_tiles_window equals mm_deconv_tiles at rows and columns [lo, hi) of every plane, bit for bit."""
    from mermake_amd import _lib as L
    from mermake_amd.deconvolver import Deconvolver, prepare_stack, tile_geometry, tile_table
    rng = np.random.default_rng(hi)
    img = rng.integers(100, 4000, size=shape).astype(np.uint16)
    psf = rng.random((shape[0], 9, 9)).astype(np.float32)
    d = Deconvolver(psf, shape, tile_size=ts, zpad=2, overlap=ov, beta=1e-3)
    x = prepare_stack(T(img, dev))
    geom = tile_geometry(tuple(x.shape), ts, ov, 1)
    table = tile_table(geom, x, None)
    n, Tt = len(geom), d._plan.T
    full = d.run_tiles(table, n, n, None).clone()
    spec = torch.empty(n * d._plan.spec_bytes, dtype=torch.uint8, device=dev)
    res = torch.full_like(full, 7.0)
    L.call("mm_deconv_tiles_window", d._plan.h, L.ptr(table), n, L.ptr(d.psf_fft), 1, n, L.ptr(spec), L.ptr(res),
           lo, hi, L.stream_ptr(dev))
    a = res[:, :, lo:hi, lo:hi].cpu().numpy()
    b = full[:, :, lo:hi, lo:hi].cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    with pytest.raises(L.MermakeLibraryError):
        L.call("mm_deconv_tiles_window", d._plan.h, L.ptr(table), n, L.ptr(d.psf_fft), 1, n, L.ptr(spec),
               L.ptr(res), 0, Tt + 1, L.stream_ptr(dev))


def test_multi_psf_pairing(dev):
    """dict of PSFs: tile i uses PSF i, zip truncates to the shorter list (deconvolver.py:110-116)."""
    from mermake_amd.deconvolver import Deconvolver
    rng = np.random.default_rng(7)
    shape = (6, 40, 40)
    img = rng.integers(0, 3000, size=shape).astype(np.uint16)
    psfs = {k: rng.random((3, 5, 5)).astype(np.float32) for k in ("a", "b", "c")}
    d = Deconvolver(psfs, shape, tile_size=20, zpad=2, overlap=4, beta=1e-3)
    assert len(d.psf_fft) == 3
    got = list(d.tile_wise(T(img, dev)))
    W = odec.wiener(psfs, shape[0], 20, 2, 4, 1e-3, mode="fp64")
    want = list(odec.tile_wise(img, W, None, None, 20, 4, 2))
    assert len(got) == len(want) == 3
    for (_, _, gt, _), (_, _, wt, _) in zip(got, want):
        assert _rel_l2(gt.cpu().numpy(), wt) < 1e-5


@pytest.mark.parametrize("zpad", [2, 5])  # 5 = sz - 1: the length-80 z pass with one PSF per tile
def test_keyed_psf_pairing_equals_zip_of_assigned_psfs(dev, zpad):
    """Keyed pairing (every tile takes the PSF whose (z, x, y) key is nearest its centre) gives
    the zip result of the per-tile PSF list it assigns: tiles centred at (10, 10), (10, 30),
    (30, 10), (30, 30) take keys 2, 0, 1 and 0 (a tie between 0 and 1 goes to the first)."""
    from mermake_amd.deconvolver import Deconvolver
    rng = np.random.default_rng(8)
    shape = (6, 40, 40)
    img = rng.integers(0, 3000, size=shape).astype(np.uint16)
    p = [rng.random((3, 5, 5)).astype(np.float32) for _ in range(3)]
    keyed = {(0, 10, 30): p[0], (0, 30, 10): p[1], (0, 10, 10): p[2]}
    d = Deconvolver(keyed, shape, tile_size=20, zpad=zpad, overlap=4, beta=1e-3, psf_mode="keyed")
    assert [t["psf"] for t in d.geometry(shape)] == [2, 0, 1, 0]
    got = list(d.tile_wise(T(img, dev)))
    assigned = {f"t{k}": p[i] for k, i in enumerate((2, 0, 1, 0))}
    W = odec.wiener(assigned, shape[0], 20, zpad, 4, 1e-3, mode="fp64")
    want = list(odec.tile_wise(img, W, None, None, 20, 4, zpad))
    assert len(got) == len(want) == 4
    for (_, _, gt, _), (_, _, wt, _) in zip(got, want):
        assert _rel_l2(gt.cpu().numpy(), wt) < 1e-5


@pytest.mark.parametrize("blur_radius,shape,ts,ov", [(50, (12, 300, 260), 100, 40), (None, (10, 230, 230), 100, 20),
                                                    (30, (8, 160, 150), 60, 21), (50, (8, 300, 290), 146, 40)])
def test_apply_window_equals_stitched_tiles(dev, blur_radius, shape, ts, ov):
    """``apply`` computes each tile only at the rows and columns [ov, T - ov) its stitch reads;
    the result equals the whole tiles of ``tile_wise`` stitched as ``deconvolver.py:179-183``,
    bit for bit."""
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.synth import gaussian_psf
    rng = np.random.default_rng(9)
    img = rng.integers(100, 4000, size=shape).astype(np.uint16)
    flat = (0.8 + 0.4 * rng.random(shape[1:])).astype(np.float32)
    d = Deconvolver(gaussian_psf((9, 21, 21), 1.5, 1.2), shape, tile_size=ts, zpad=(shape[0] - 1) // 2, overlap=ov,
                    beta=0.01)
    got = d.apply(T(img, dev), flat_field=T(flat, dev), blur_radius=blur_radius).cpu().numpy()
    want = np.zeros(shape, np.float32)
    for x, y, t, _ in d.tile_wise(T(img, dev), flat_field=T(flat, dev), blur_radius=blur_radius):
        t = t.cpu().numpy()
        xe, ye = t.shape[1], t.shape[2]
        want[:, x:x + xe - 2 * ov, y:y + ye - 2 * ov] = t[:, ov:xe - ov, ov:ye - ov]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_multi_psf_length80_zpass_40_planes(dev):
    """40 planes, zpad = 39, one PSF per tile (the grid form of the length-80 z pass, k_tile_zfft_own<40>)."""
    from mermake_amd.deconvolver import Deconvolver
    rng = np.random.default_rng(10)
    shape = (40, 60, 60)
    img = rng.integers(100, 4000, size=shape).astype(np.uint16)
    psfs = {k: rng.random((5, 7, 7)).astype(np.float32) + np.float32(0.1) for k in ("a", "b", "c", "d")}
    d = Deconvolver(psfs, shape, tile_size=30, zpad=39, overlap=6, beta=1e-3)
    got = list(d.tile_wise(T(img, dev)))
    W = odec.wiener(psfs, shape[0], 30, 39, 6, 1e-3, mode="fp64")
    want = list(odec.tile_wise(img, W, None, None, 30, 6, 39))
    assert len(got) == len(want) == 4
    for (_, _, gt, _), (_, _, wt, _) in zip(got, want):
        assert _rel_l2(gt.cpu().numpy(), wt) < 1e-5


def test_wiener_closed_form(dev):
    """Delta PSF: W = 1/(1 + beta |L|^2) / (Zp T T); L(0) = 1 so the DC gain is 1/(1+beta)."""
    from mermake_amd.deconvolver import Deconvolver
    sz, ts, ov, zpad, beta = 6, 20, 3, 2, 1e-2
    psf = np.zeros((sz, ts, ts), np.float32)
    psf[sz // 2, ts // 2, ts // 2] = 1
    d = Deconvolver(psf, (sz, 40, 40), tile_size=ts, zpad=zpad, overlap=ov, beta=beta)
    Zp, Tt = sz + 2 * zpad, ts + 2 * ov
    ops = d.psf_fft[0].cpu().numpy().astype(np.complex128)
    assert ops.shape[0] == 2 * Zp       # stacks of <= 48 planes: w = unnormalised ifft_z(W), then W itself
    L2 = odec.laplacian_power((Zp, Tt, Tt))[:, :, :Tt // 2 + 1]
    want = 1.0 / (1.0 + beta * L2) / (Zp * Tt * Tt)
    for W in (np.fft.fft(ops[:Zp], axis=0) / Zp, ops[Zp:]):
        assert np.abs(W - want).max() / want.max() < 1e-6
        assert abs(W[0, 0, 0].real * Zp * Tt * Tt - 1 / (1 + beta)) < 1e-6


def test_fixture_golden_tiles(dev):
    """The reference's golden float16 deconvolution of the real fixture
    (tests/test_deconvolver.py:422-434) on three committed tile windows."""
    from fixtures import OV, TS, fixture_stack, golden_accept, psf_750
    from mermake_amd.deconvolver import Deconvolver
    stack, wins, gz = fixture_stack()
    psf = psf_750()
    d = Deconvolver(psf, stack.shape, tile_size=TS, zpad=stack.shape[0], overlap=OV, beta=0.001)
    out = d.apply(T(stack, dev)).cpu().numpy()
    W = odec.wiener(psf, stack.shape[0], TS, stack.shape[0], OV, 0.001, mode="fp64")
    only = set(wins)
    for x, y, t, _ in odec.tile_wise(stack, W, None, None, TS, OV, stack.shape[0], only=only):
        mine = out[:, x:x + TS, y:y + TS]
        assert _rel_l2(mine, t[:, OV:OV + TS, OV:OV + TS]) < 1e-5, (x, y)
        exact, worst = golden_accept(mine[list(gz)], wins[(x, y)][2])
        assert exact > 0.995 and worst <= 1.0, (x, y, exact, worst)


@pytest.mark.parametrize("n,offset", [(1, 0), (3, 0), (4096 * 37 + 3, 0), (4096 * 37 + 3, 1), (5000, 1),
                                      (40 * 300 * 301, 0)])
def test_rms_normalize_matches_float64_norm(dev, n, offset):
    """mm_rms_normalize (__main__.py:221-222): the divisor is the float64 norm over sqrt(N) rounded to float32,
    and every element is divided by it in float32, for 16-byte aligned arrays (float4 loads and a tail) and
    unaligned ones (element kernels)."""
    from mermake_amd import _lib as L
    rng = np.random.default_rng(n + offset)
    a = (rng.standard_normal(n) * 100).astype(np.float32)
    buf = torch.zeros(n + offset, dtype=torch.float32, device=dev)
    view = buf[offset:]
    view.copy_(torch.from_numpy(a))
    ws = torch.empty(L.lib().mm_rms_ws_bytes(), dtype=torch.uint8, device=dev)
    rms = torch.empty(1, dtype=torch.float32, device=dev)
    L.call("mm_rms_normalize", L.ptr(view), n, L.ptr(ws), L.ptr(rms), L.stream_ptr(dev))
    want_rms = np.float32(np.sqrt(np.sum(a.astype(np.float64) ** 2)) / np.sqrt(np.float64(n)))
    assert rms.item() == want_rms
    assert np.array_equal(view.cpu().numpy(), a / want_rms)
    if offset:
        assert buf[0].item() == 0.0


@pytest.mark.parametrize("shape,delta,thr", [((12, 40, 70), 5, 3.0), ((10, 33, 130), 1, 2.0), ((9, 20, 64), 2, -0.5),
                                             ((6, 6, 6), 5, 0.0), ((1, 1, 1), 1, -1.0), ((5, 30, 70), 3, 1.0),
                                             ((4, 9, 130), 4, 0.5), ((11, 11, 200), 5, 1.5)])
def test_maxima_both_signs_in_one_pass(dev, shape, delta, thr):
    """find_local_maxima_pm (mm_local_maxima_pm) gives the + and - maxima and fits of the two signed calls
    bit for bit, with the candidate sets of the oracle; a negative threshold makes voxels candidates of
    both signs."""
    from mermake_amd.maxima import find_local_maxima_pm, find_local_maxima_signed
    rng = np.random.default_rng(shape[2] + delta)
    img = (rng.standard_normal(shape) * 2).astype(np.float32)
    raw = rng.integers(0, 4000, shape).astype(np.uint16)
    kw = dict(threshold=thr, delta=delta, delta_fit=3, sigmaZ=1.0, sigmaXY=1.5)
    ti, tr = T(img, dev), T(raw, dev)
    plus, minus = find_local_maxima_pm(ti, tr, **kw)
    p1 = find_local_maxima_signed(ti, tr, sign=1.0, **kw)
    m1 = find_local_maxima_signed(ti, tr, sign=-1.0, **kw)
    for a, b in ((plus, p1), (minus, m1)):   # (the correlations of a flat neighbourhood are NaN)
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy(), equal_nan=True)
    for sgn, got in ((1.0, plus), (-1.0, minus)):
        z, x, y = omax.local_maxima(sgn * img, thr, delta)
        assert got.shape[0] == len(z)
