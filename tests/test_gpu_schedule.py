"""Schedules of the tiled deconvolution: every pass form a tile geometry admits gives the float64
oracle's tiles (relative L2 <= 1e-5), and the autotuner selects the fastest of them, caches its
choice and keeps the result."""
import itertools

import numpy as np
import pytest
import torch

from oracle import deconv as odec

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _rel_l2(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


def _case(shape, npsf, flat, seed=11):
    rng = np.random.default_rng(seed)
    img = rng.integers(100, 4000, size=shape).astype(np.uint16)
    psfs = [rng.random((5, 7, 7)).astype(np.float32) + np.float32(0.1) for _ in range(npsf)]
    fl = (0.7 + 0.6 * rng.random(shape[1:])).astype(np.float32) if flat else None
    return img, (psfs[0] if npsf == 1 else {f"p{i}": p for i, p in enumerate(psfs)}), fl


# (shape, tile_size, overlap, zpad, PSFs, flat, admitted y/x forms, admitted z forms)
GEOMS = [
    ((40, 260, 260), 200, 13, 39, 1, True, {"stockham", "prime_first", "rader"}, {"fft", "matrix", "dft80"}),
    ((40, 260, 260), 200, 13, 39, 2, False, {"stockham", "prime_first", "rader"}, {"fft", "matrix", "dft80"}),
    ((3, 600, 600), 500, 89, 3, 1, True, {"stockham", "prime_first", "rader"}, {"fft", "matrix"}),
    ((10, 330, 330), 300, 89, 9, 1, False, None, {"fft", "matrix", "dft80"}),
    ((4, 70, 70), 50, 8, 2, 2, True, {"stockham", "prime_first"}, {"fft", "matrix"}),
    ((40, 120, 120), 80, 10, 39, 1, False, {"stockham"}, {"fft", "matrix", "dft80"}),
    ((50, 140, 140), 100, 20, 3, 1, False, {"stockham"}, {"fft"}),
    ((40, 260, 260), 200, 13, 19, 1, True, {"stockham", "prime_first", "rader"}, {"fft", "matrix"}),
    ((40, 260, 260), 200, 13, 19, 2, False, {"stockham", "prime_first", "rader"}, {"fft", "matrix"}),
]


@pytest.mark.parametrize("shape,ts,ov,zpad,npsf,flat,xy_forms,z_forms", GEOMS)
def test_every_schedule_vs_oracle(dev, shape, ts, ov, zpad, npsf, flat, xy_forms, z_forms):
    """T = 226 = 113 * 2 and T = 678 = 113 * 6 (Rader, prime-first, Stockham), T = 478 = 239 * 2,
    T = 66 = 11 * 6, T = 100 and T = 140 (Stockham only); 40 planes with zpad = sz - 1 (all three z
    forms), 3 and 4 planes (no length-80 pass), 50 planes (above the z-kernel matrix's 48); 40 planes
    with zpad 19 (Zp 78, the DAPI setting: the z FFT form runs its length-78 transforms in registers),
    one PSF and one per tile."""
    from mermake_amd.deconvolver import Deconvolver
    img, psfs, fl = _case(shape, npsf, flat)
    d = Deconvolver(psfs, shape, tile_size=ts, zpad=zpad, overlap=ov, beta=1e-3)
    s = d.schedule()
    if xy_forms is not None:
        assert set(s["xy_forms"]) == xy_forms
    assert set(s["z_forms"]) == z_forms
    assert "rader" in s["xy_forms"] or shape[1] < 200
    W = odec.wiener(psfs, shape[0], ts, zpad, ov, 1e-3, mode="fp64")
    want = [t for _, _, t, _ in odec.tile_wise(img, W, fl, None, ts, ov, zpad, mode="fp64")]
    flat_d = None if fl is None else T(fl, dev)
    img_d = T(img, dev)
    for xy, z in itertools.product(s["xy_forms"], s["z_forms"]):
        d.set_schedule(xy, z)
        now = d.schedule()
        assert (now["xy"], now["z"]) == (xy, z)
        got = [t.cpu().numpy() for _, _, t, _ in d.tile_wise(img_d, flat_field=flat_d)]
        assert len(got) == len(want)
        for i, (g, w) in enumerate(zip(got, want)):
            assert _rel_l2(g, w) < 1e-5, (xy, z, i)


def test_schedule_rejects_forms_the_geometry_does_not_admit(dev):
    from mermake_amd._lib import MermakeLibraryError
    from mermake_amd.deconvolver import Deconvolver
    shape = (4, 70, 70)
    img, psf, _ = _case(shape, 1, False)
    d = Deconvolver(psf, shape, tile_size=50, zpad=2, overlap=8, beta=1e-3)
    before = d.schedule()
    with pytest.raises(MermakeLibraryError):
        d.set_schedule("rader", before["z"])          # T = 66 has no Rader prime
    with pytest.raises(MermakeLibraryError):
        d.set_schedule(before["xy"], "dft80")         # zpad != sz - 1
    with pytest.raises(ValueError):
        d.set_schedule("radix9", before["z"])
    assert d.schedule() == before
    W = odec.wiener(psf, shape[0], 50, 2, 8, 1e-3, mode="fp64")
    want = [t for _, _, t, _ in odec.tile_wise(img, W, None, None, 50, 8, 2, mode="fp64")]
    got = [t.cpu().numpy() for _, _, t, _ in d.tile_wise(T(img, dev))]
    assert all(_rel_l2(g, w) < 1e-5 for g, w in zip(got, want))


def test_autotune_selects_the_fastest_and_caches_it(dev, tmp_path, monkeypatch):
    """The tuner times every admitted (y/x, z) pair, keeps the least time, then the tiles per launch;
    a second Deconvolver of the same geometry reads the choice from the cache file, and the tuned
    Deconvolver's tiles equal the oracle's."""
    from mermake_amd.deconvolver import Deconvolver
    monkeypatch.setenv("MERMAKE_AMD_TUNE_CACHE", str(tmp_path / "schedule.json"))
    shape = (40, 260, 260)
    img, psf, fl = _case(shape, 1, True)
    kw = dict(tile_size=200, zpad=39, overlap=13, beta=1e-3)
    d = Deconvolver(psf, shape, autotune={"reps": 2}, **kw)
    t = d.tuning
    s = d.schedule()
    assert not t["cached"] and t["tiles"] == 4
    assert set(t["ms_forms"]) == {f"{a}+{b}" for a in s["xy_forms"] for b in s["z_forms"]}
    assert all(v > 0 for v in t["ms_forms"].values())
    assert min(t["ms_forms"], key=t["ms_forms"].get) == f"{t['xy']}+{t['z']}" == f"{s['xy']}+{s['z']}"
    assert set(t["ms_batch_tiles"]) == {"4", "2", "1"}
    fastest = min(t["ms_batch_tiles"].values())
    B = max(int(b) for b, ms in t["ms_batch_tiles"].items() if ms <= 1.02 * fastest)
    assert d.batch_tiles == (None if B == 4 else B) == s["batch_tiles"]
    assert (tmp_path / "schedule.json").exists()
    d2 = Deconvolver(psf, shape, autotune=True, **kw)
    assert d2.tuning["cached"] and (d2.tuning["xy"], d2.tuning["z"]) == (t["xy"], t["z"])
    assert d2.schedule()["xy"] == t["xy"] and d2.batch_tiles == d.batch_tiles
    W = odec.wiener(psf, shape[0], 200, 39, 13, 1e-3, mode="fp64")
    want = [t_ for _, _, t_, _ in odec.tile_wise(img, W, fl, None, 200, 13, 39, mode="fp64")]
    got = [t_.cpu().numpy() for _, _, t_, _ in d2.tile_wise(T(img, dev), flat_field=T(fl, dev))]
    assert all(_rel_l2(g, w) < 1e-5 for g, w in zip(got, want))


def test_hyb_pipeline_autotuned_equals_oracle_spots(dev, tmp_path, monkeypatch):
    """``HybPipeline.from_settings(autotune=...)`` tunes on the pipeline's deconvolution window and
    finds the spots of the default schedule within the end-to-end tolerance (same candidates after the
    same deconvolution up to float32 rounding)."""
    from mermake_amd.pipeline import HybPipeline
    from mermake_amd.synth import gaussian_psf, stack_np
    monkeypatch.setenv("MERMAKE_AMD_TUNE_CACHE", str(tmp_path / "schedule.json"))
    shape = (40, 300, 300)
    raw, flat = stack_np(5, shape, 400, flat=True)
    psf = gaussian_psf((11, 31, 31), 1.5, 1.2)
    hyb = dict(blur_radius=30, threshold=3600, delta=1, delta_fit=3, sigmaZ=1, sigmaXY=1.5)
    base = HybPipeline.from_settings(psf, shape, tile_size=200, overlap=13, beta=1e-4, **hyb)
    tuned = HybPipeline.from_settings(psf, shape, tile_size=200, overlap=13, beta=1e-4, autotune={"reps": 1}, **hyb)
    assert tuned.d.tuning is not None and not tuned.d.tuning["cached"]
    a = base(T(raw, dev), T(flat, dev)).cpu().numpy()
    b = tuned(T(raw, dev), T(flat, dev)).cpu().numpy()
    assert len(a) > 50 and abs(len(a) - len(b)) <= max(2, len(a) // 200)
    ka = {tuple(r) for r in np.round(a[:, :3]).astype(int)}
    kb = {tuple(r) for r in np.round(b[:, :3]).astype(int)}
    assert len(ka & kb) >= 0.99 * len(ka)
