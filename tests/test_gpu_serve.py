"""Batched multi-FOV execution and the continuous-batching request loop: each FOV's spots are
the same alone and in the mix (tiles of uneven FOV sizes share launches)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HYB = dict(blur_radius=30, threshold=3600, delta=1, delta_fit=3, sigmaZ=1, sigmaXY=1.5)
SHAPES = [(12, 260, 300), (12, 230, 230), (12, 140, 190), (12, 300, 120)]


@pytest.fixture(scope="module")
def fovs():
    from mermake_amd.synth import gaussian_psf, stack_np
    out = []
    for i, shape in enumerate(SHAPES):
        chans = []
        flat = None
        for c in range(2):
            raw, f = stack_np(100 + 10 * i + c, shape, 300, flat=True)
            chans.append(raw)
            flat = f
        out.append((chans, flat))
    psf = gaussian_psf((11, 31, 31), 1.5, 1.2)
    return out, psf


@pytest.fixture(scope="module")
def pipe(fovs, dev):
    from mermake_amd.pipeline import HybPipeline
    _, psf = fovs
    return HybPipeline.from_settings(psf, SHAPES[0], tile_size=100, overlap=20, beta=1e-4, **HYB)


@pytest.fixture(scope="module")
def alone(fovs, pipe, dev):
    """Each channel through the single-stack pipeline."""
    data, _ = fovs
    res = []
    for chans, flat in data:
        fl = torch.from_numpy(flat).to(dev)
        res.append([pipe(torch.from_numpy(c).to(dev), fl).cpu().numpy() for c in chans])
    assert sum(len(x) for r in res for x in r) > 100
    return res


def test_run_many_uneven_fovs_equals_alone(fovs, pipe, alone, dev):
    data, _ = fovs
    stacks, flats, where = [], [], []
    for i, (chans, flat) in enumerate(data):
        fl = torch.from_numpy(flat).to(dev)
        for c, ch in enumerate(chans):
            stacks.append(torch.from_numpy(ch).to(dev))
            flats.append(fl)
            where.append((i, c))
    got = pipe.run_many(stacks, flats)
    for (i, c), g in zip(where, got):
        assert np.array_equal(g.cpu().numpy(), alone[i][c]), (i, c)


def test_run_many_host_stacks_equal_alone(fovs, pipe, alone, dev):
    """Host NumPy stacks (each copied to the device inside ``run_many``, with the flats cast from
    float64) give the spots of each stack alone: the copies stay alive until every batch reading them
    has finished, while later copies are allocated."""
    data, _ = fovs
    stacks, flats, where = [], [], []
    for i, (chans, flat) in enumerate(data):
        for c, ch in enumerate(chans):
            stacks.append(np.ascontiguousarray(ch))
            flats.append(flat.astype(np.float64))
            where.append((i, c))
    got = pipe.run_many(stacks, flats)
    for (i, c), g in zip(where, got):
        assert np.array_equal(g.cpu().numpy(), alone[i][c]), (i, c)


@pytest.mark.parametrize("host_inputs,batch_tiles,budget_fovs", [(False, 5, None), (True, 7, 1.5), (True, 36, None)])
def test_server_mix_equals_alone(fovs, pipe, alone, dev, host_inputs, batch_tiles, budget_fovs):
    """Requests arriving while others run are admitted as memory frees up and batched together;
    every request's spots equal its channels run alone."""
    from mermake_amd.serve import FovRequest, HybServer
    data, _ = fovs
    reqs = []
    arrivals = [0.0, 0.0, 0.005, 0.02]
    for i, (chans, flat) in enumerate(data):
        if host_inputs:
            ch = [torch.from_numpy(c).pin_memory() for c in chans]
            fl = torch.from_numpy(flat).pin_memory()
        else:
            ch = [torch.from_numpy(c).to(dev) for c in chans]
            fl = torch.from_numpy(flat).to(dev)
        reqs.append(FovRequest(i, ch, fl, arrival=arrivals[i]))
    budget = None
    if budget_fovs is not None:
        per = max(sum(c.nbytes for c in chans) + flat.nbytes for chans, flat in data)
        budget = int(budget_fovs * per)
    srv = HybServer(pipe, batch_tiles=batch_tiles, slots=2, hbm_budget=budget)
    done = srv.run(reqs)
    assert len(done) == len(reqs)
    for r in done:
        assert r.t_done is not None and r.latency >= 0
        for c, s in enumerate(r.spots):
            assert np.array_equal(s, alone[r.rid][c]), (r.rid, c)
    st = srv.stats
    assert st["requests"] == 4 and st["p99_latency_s"] >= st["p50_latency_s"] > 0
    ntiles = sum(2 * -(-sh[1] // 100) * -(-sh[2] // 100) for sh in SHAPES)
    assert st["tiles"] == ntiles and st["launches"] >= -(-ntiles // batch_tiles)
    if budget_fovs is not None:
        # with room for one request at a time, admissions follow completions
        order = sorted(done, key=lambda r: r.t_admit)
        for a, b in zip(order, order[1:]):
            assert b.t_admit >= a.t_done - 1e-9


def test_server_rejects_oversized_request(pipe, fovs):
    from mermake_amd.serve import FovRequest, HybServer
    data, _ = fovs
    chans, flat = data[0]
    srv = HybServer(pipe, batch_tiles=4, slots=1, hbm_budget=1024)
    with pytest.raises(ValueError):
        srv.run([FovRequest(0, [torch.from_numpy(c) for c in chans], torch.from_numpy(flat))])


def test_server_keyed_multi_psf_equals_alone(fovs, dev):
    """Several PSFs keyed by position (the served stream's pairing): every tile of every FOV
    size takes its nearest PSF, in the mix as alone."""
    from mermake_amd.pipeline import HybPipeline
    from mermake_amd.serve import FovRequest, HybServer
    from mermake_amd.synth import gaussian_psf
    data, _ = fovs
    psfs = {(0, x, y): gaussian_psf((11, 31, 31), 1.2 + 0.2 * i, 1.0 + 0.15 * i)
            for i, (x, y) in enumerate(((75, 75), (75, 225), (225, 75), (225, 225)))}
    pk = HybPipeline.from_settings(psfs, SHAPES[0], tile_size=100, overlap=20, beta=1e-4, psf_mode="keyed", **HYB)
    used = {t["psf"] for shape in SHAPES for t in pk.d.geometry(shape)}
    assert used == {0, 1, 2, 3}
    alone = []
    for chans, flat in data:
        fl = torch.from_numpy(flat).to(dev)
        alone.append([pk(torch.from_numpy(c).to(dev), fl).cpu().numpy() for c in chans])
    srv = HybServer(pk, batch_tiles=7, slots=2)
    reqs = [FovRequest(i, [torch.from_numpy(c).to(dev) for c in chans], torch.from_numpy(flat).to(dev),
                       arrival=0.0) for i, (chans, flat) in enumerate(data)]
    done = srv.run(reqs)
    for r in done:
        for c, sp in enumerate(r.spots):
            assert np.array_equal(sp, alone[r.rid][c]), (r.rid, c)


def test_server_squared_u8_requests_equal_uint16(fovs, pipe, dev):
    """FOVs sent as pinned ``SquaredU8`` frames (half the bytes) give the spots of their uint16
    squares, bit for bit, mixed with uint16 requests."""
    from mermake_amd.deconvolver import SquaredU8
    from mermake_amd.serve import FovRequest, HybServer
    data, _ = fovs
    reqs, want = [], []
    for i, (chans, flat) in enumerate(data):
        u8 = [np.minimum(np.sqrt(c), 255).astype(np.uint8) for c in chans]
        sq = [u.astype(np.uint16) ** 2 for u in u8]
        fl = torch.from_numpy(flat)
        want.append([pipe(torch.from_numpy(c).to(dev), fl.to(dev)).cpu().numpy() for c in sq])
        ch = [SquaredU8.of(torch.from_numpy(u)).pin_memory() if i % 2 == 0 else torch.from_numpy(c).pin_memory()
              for u, c in zip(u8, sq)]
        reqs.append(FovRequest(i, ch, fl.pin_memory(), arrival=0.0))
    assert sum(len(x) for w in want for x in w) > 20
    done = HybServer(pipe, batch_tiles=7, slots=2).run(reqs)
    for r in done:
        for c, s in enumerate(r.spots):
            assert np.array_equal(s, want[r.rid][c]), (r.rid, c)
