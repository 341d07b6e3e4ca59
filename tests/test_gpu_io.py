"""GPU tests of the host feed: ``load_flats`` (HIP box blur, median, division) against the
oracle, and the pinned double-buffered ``FovFeeder``."""
import numpy as np
import pytest
import torch

from mermake_amd import io as mio
from oracle import io as oio

from fixtures import make_fov

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(256, 290), (255, 291)])  # even and odd pixel counts (median of 2 / of 1)
def test_load_flats_bitexact(dev, tmp_path, shape):
    rng = np.random.default_rng(5)
    ims = [rng.random((300, 330), dtype=np.float32) * 1000 + 200,
           rng.integers(100, 60000, (280, 300)).astype(np.uint16)]
    for i, im in enumerate(ims):
        np.savez(tmp_path / f"Scope0_{i}.npz", im=im)
    got = mio.load_flats(str(tmp_path / "Scope0"), shape, device=dev)
    assert got.shape == (2,) + shape and got.is_cuda
    want = np.stack([oio.load_flat(im, shape) for im in ims])
    np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_load_flats_constant_is_ones(dev, tmp_path):
    np.savez(tmp_path / "flat_a.npz", im=np.full((64, 70), 3.5, np.float32))
    got = mio.load_flats(str(tmp_path / "flat_"), (50, 60), device=dev).cpu().numpy()
    np.testing.assert_array_equal(got, np.ones((1, 50, 60), np.float32))
    with pytest.raises(FileNotFoundError):
        mio.load_flats(str(tmp_path / "none_"), (50, 60), device=dev)


def test_fov_feeder_stages_fovs_to_device(dev, tmp_path):
    """Three FOVs of different depths through the two pinned staging buffers (the third reuses
    the first buffer, grown): every channel on the device equals ``read_im``'s."""
    rng = np.random.default_rng(6)
    paths, want = [], []
    for i, n in enumerate((10, 13, 7)):
        frames = rng.integers(0, 256, (n, 40, 48), dtype=np.uint8)
        p = make_fov(str(tmp_path / "H1_set2"), f"{i:03d}", frames, pos=f"{i}.5,-{i}.25")
        paths.append(p)
        want.append(mio.read_im(p))
    print("feeder FOVs", [w.shape for w in want], flush=True)
    with mio.FovFeeder(paths, device=dev) as feed:
        got = [(p, [c.cpu().numpy() for c in chans], x, y) for p, chans, x, y in feed]
    assert [g[0] for g in got] == paths
    for i, (p, chans, x, y) in enumerate(got):
        assert (x, y) == (i + 0.5, -(i + 0.25))
        assert len(chans) == 3
        for c in range(3):
            np.testing.assert_array_equal(chans[c], want[i][c])


def test_fov_feeder_u8_hands_out_squared_u8(dev, tmp_path):
    """``u8=True``: the channels are ``SquaredU8`` views of the uint8 frames, whose squares are
    ``read_im``'s channels."""
    from mermake_amd.deconvolver import SquaredU8
    rng = np.random.default_rng(8)
    frames = rng.integers(0, 256, (1 + 3 * 5, 24, 40), dtype=np.uint8)
    p = make_fov(str(tmp_path / "H2_set1"), "004", frames)
    want = mio.read_im(p)
    with mio.FovFeeder([p], device=dev, u8=True) as feed:
        (_, chans, _, _), = list(feed)
    assert feed.h2d_bytes == want.size  # one byte a pixel
    for c in range(3):
        assert isinstance(chans[c], SquaredU8) and chans[c].dtype == torch.uint8
        np.testing.assert_array_equal(chans[c].cpu().numpy().astype(np.uint16) ** 2, want[c])


def test_fov_feeder_reports_unreadable_fov(dev, tmp_path):
    frames = np.zeros((4, 8, 8), np.uint8)
    good = make_fov(str(tmp_path / "H1_set1"), "001", frames)
    feed = mio.FovFeeder([good, str(tmp_path / "H1_set1" / "Conv_zscan__009.zarr")], device=dev)
    try:
        p, chans, _, _ = next(feed)
        assert p == good and chans[0].shape == (1, 8, 8)
        with pytest.raises(RuntimeError):
            next(feed)
    finally:
        feed.close()


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 1001, 4096 * 1024 + 3])
def test_median_matches_numpy(dev, n):
    """``mm_median`` (radix selection) equals numpy.median in float32: odd and even counts, duplicates,
    negative values and zeros of both signs."""
    from mermake_amd.io import _median
    rng = np.random.default_rng(n)
    a = (rng.standard_normal(n) * 100).astype(np.float32)
    a[: n // 3] = np.round(a[: n // 3])            # many duplicates
    if n > 4:
        a[:2] = [0.0, -0.0]
    got = float(_median(torch.from_numpy(a).to(dev)).item())
    want = np.median(a)
    assert got == np.float32(want), (got, want)
    if n > 2:
        b = a.copy()
        b[n // 2] = np.nan
        assert np.isnan(float(_median(torch.from_numpy(b).to(dev)).item()))
