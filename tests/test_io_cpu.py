"""Host feed and output formats (``mermake_amd.io``) on the CPU: the zarr v2 reader, ``read_im``,
the reference's output names and ``.npz`` layout.  ``load_flats`` and the device feeder are in
``test_gpu_io.py``."""
import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

from mermake_amd import io as mio
from oracle import io as oio

from fixtures import fixture_windows, make_fov, write_zarr

REF = "/root/reference/tests/examples/single"


@pytest.mark.parametrize("compressor,order,sep", [("bz2", "C", "."), ("zlib", "F", "/"), (None, "C", ".")])
def test_read_zarr_round_trip(tmp_path, compressor, order, sep):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 60000, (7, 13, 10), dtype=np.uint16)
    p = str(tmp_path / "arr")
    write_zarr(p, a, (3, 5, 4), compressor, order, sep)
    np.testing.assert_array_equal(mio.read_zarr(p), a)
    np.testing.assert_array_equal(mio.read_zarr(p, first=1), a[1:])
    np.testing.assert_array_equal(mio.read_zarr(p, first=4, workers=1), a[4:])


def test_read_zarr_missing_chunk_reads_fill_value(tmp_path):
    a = np.arange(4 * 6, dtype=np.float32).reshape(4, 6)
    p = str(tmp_path / "arr")
    write_zarr(p, a, (2, 3), None, fill_value=-1.5, skip={(1, 0)})
    exp = a.copy()
    exp[2:4, 0:3] = -1.5
    np.testing.assert_array_equal(mio.read_zarr(p), exp)


def test_read_zarr_rejects_unknown_codec(tmp_path):
    p = str(tmp_path / "arr")
    write_zarr(p, np.zeros((2, 2), np.uint8), (2, 2), None)
    meta = json.load(open(os.path.join(p, ".zarray")))
    meta["compressor"] = {"id": "blosc"}
    json.dump(meta, open(os.path.join(p, ".zarray"), "w"))
    with pytest.raises(NotImplementedError):
        mio.read_zarr(p)


def test_read_im_deinterleaves_and_squares(tmp_path):
    rng = np.random.default_rng(2)
    frames = rng.integers(0, 256, (1 + 3 * 4 + 2, 9, 11), dtype=np.uint8)  # 2 frames beyond the last full z
    path = make_fov(str(tmp_path / "H0_set7"), "031", frames)
    im, x, y = mio.read_im(path, return_pos=True)
    assert im.shape == (3, 4, 9, 11) and im.dtype == np.uint16
    assert (x, y) == (-4351.10, 5862.40)
    for c in range(3):
        for k in range(4):  # channel-major index identity of io.py:140-142
            np.testing.assert_array_equal(im[c, k], frames[1 + 3 * k + c].astype(np.uint16) ** 2)
    np.testing.assert_array_equal(im, oio.channels_from_frames(frames, 3))
    assert mio.get_ifov(path) == 31 and mio.get_iset(path) == 7


def test_read_im_unsquared_uint8(tmp_path):
    rng = np.random.default_rng(9)
    frames = rng.integers(0, 256, (1 + 3 * 2, 6, 7), dtype=np.uint8)
    path = make_fov(str(tmp_path / "H3_set1"), "005", frames)
    raw = mio.read_im(path, square=False)
    assert raw.dtype == np.uint8 and raw.shape == (3, 2, 6, 7)
    np.testing.assert_array_equal(raw.astype(np.uint16) ** 2, mio.read_im(path))


def test_read_im_without_xml(tmp_path):
    frames = np.arange(5 * 4 * 4, dtype=np.uint8).reshape(5, 4, 4)
    path = make_fov(str(tmp_path / "d"), "002", frames)
    os.remove(os.path.splitext(path)[0] + ".xml")
    im = mio.read_im(path)
    np.testing.assert_array_equal(im, frames[1:].astype(np.uint16) ** 2)
    with pytest.raises(ValueError):
        mio.read_im(path, return_pos=True)


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's fixture FOV is not mounted")
def test_read_im_reference_fixture_matches_golden_windows():
    """The reference's fixture FOV (108 bz2 frames of 1200^2, 3 channels) read by ``read_im``:
    channel 0 holds the squared input windows of the committed golden tiles."""
    im = mio.read_im(os.path.join(REF, "H1_TEST_set1/Conv_zscan__020.zarr"))
    assert im.shape == (3, 35, 1200, 1200) and im.dtype == np.uint16
    wins, _, shape = fixture_windows()
    assert im[0].shape == shape
    for (x0, x1, y0, y1), w, _ in wins.values():
        np.testing.assert_array_equal(im[0][:, x0:x1, y0:y1], w)


def test_output_names_match_the_reference_files():
    """File names of the reference's own outputs for its fixture FOV
    (``tests/examples/single/out``)."""
    path = "/data/H1_TEST_set1/Conv_zscan__020.zarr"
    names = mio.OutputNames()
    assert names.get_name(path, 0) == "Conv_zscan__020--H1_TEST_set1--col0__Xhfits.npz"
    assert names.get_name(path, -1) == "Conv_zscan__020--H1_TEST_set1--dapiFeatures.npz"
    assert names.get_name(path) == "drift_Conv_zscan__020--_set1.pkl"
    assert mio.get_ifov(path) == 20 and mio.get_iset(path) == 1
    # the hyb round of a file is the first number of its folder (io.py:413, hsorted)
    assert mio.get_ih(os.path.dirname("/data/H12_x/Conv_zscan__020.zarr")) == 12
    assert mio.get_ih("/data/nodigits") == 10 ** 100
    with pytest.raises(ValueError):
        mio.get_ifov("/data/x_set1/none.zarr")


def test_namespace_to_array_rows():
    ns = SimpleNamespace(paths=SimpleNamespace(codebook="none", psf_file="p.npy"),
                         hybs={"threshold": 3600, "blur": {"radius": 30}}, version=1)
    assert mio.namespace_to_array(ns) == [("paths", "codebook", "none"), ("paths", "psf_file", "p.npy"),
                                          ("hybs", "threshold", "3600"), ("hybs.blur", "radius", "30"),
                                          ("", "version", "1")]


def test_save_xfits_layout(tmp_path):
    """The reference's result files hold ``Xh`` (N, 8) float32 or ``Xh_plus`` / ``Xh_minus``,
    a 0-d string ``version`` and an (n, 3) string ``args``."""
    rows = mio.namespace_to_array(SimpleNamespace(paths=SimpleNamespace(codebook="none"), hybs={"delta": 1}))
    Xh = np.random.default_rng(3).random((5, 8), dtype=np.float32)
    f = str(tmp_path / "a--b--col0__Xhfits.npz")
    assert mio.save_xfits(f, {"Xh": Xh}, rows, version="0.0.69")
    assert not mio.save_xfits(f, {"Xh": Xh[:1]}, rows)  # an existing file is kept
    with np.load(f) as d:
        assert sorted(d.files) == ["Xh", "args", "version"]
        np.testing.assert_array_equal(d["Xh"], Xh)
        assert d["version"].dtype.kind == "U" and d["version"].shape == () and str(d["version"]) == "0.0.69"
        assert d["args"].dtype.kind == "U" and d["args"].shape == (2, 3)
        assert d["args"][1].tolist() == ["hybs", "delta", "1"]
    f2 = str(tmp_path / "a--b--dapiFeatures.npz")
    mio.save_xfits(f2, {"Xh_plus": Xh, "Xh_minus": Xh[:2]}, rows)
    with np.load(f2) as d:
        assert sorted(d.files) == ["Xh_minus", "Xh_plus", "args", "version"]


def test_oracle_load_flat_closed_forms():
    """A constant flat normalises to ones; the median is that of a full sort."""
    c = np.full((40, 50), 7.25, np.float32)
    np.testing.assert_array_equal(oio.load_flat(c, (30, 40)), np.ones((30, 40), np.float32))
    rng = np.random.default_rng(4)
    im = rng.random((33, 45), dtype=np.float32) + 0.5
    from oracle import blur as ob
    b = ob.box(im, (20, 20), axes=(-1, -2))
    s = np.sort(b.ravel())
    n = s.size
    med = s[n // 2] if n % 2 else np.float32((s[n // 2 - 1] + s[n // 2]) / np.float32(2))
    np.testing.assert_array_equal(oio.load_flat(im), (b / med).astype(np.float32))
