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


def test_graph_replays_equal_eager(fovs, pipe, alone, dev):
    """``HybPipeline(graph=True)`` runs the first batch of a shape eagerly, captures the batch's kernel
    chain as a hipGraph and replays it for the next batches of that shape: every replay gives the
    eager spots, with other eager work and synchronisations between the replays; batches of another
    shape (uneven FOVs in one ``run_many``) get their own graph."""
    from mermake_amd.pipeline import HybPipeline
    data, _ = fovs
    gp = HybPipeline(pipe.d, graph=True, **HYB)
    chans, flat = data[0]
    fl = torch.from_numpy(flat).to(dev)
    for rep in range(3):
        for c, ch in enumerate(chans):
            got = gp(torch.from_numpy(ch).to(dev), fl)
            assert np.array_equal(got.cpu().numpy(), alone[0][c]), (rep, c)
            x = torch.randn(1 << 20, device=dev)
            assert float((x * x).sum()) > 0
            torch.cuda.synchronize()
    assert len(gp._graphs) == 1
    stacks, flats, where = [], [], []
    for i, (chs, f) in enumerate(data):
        fd = torch.from_numpy(f).to(dev)
        for c, ch in enumerate(chs):
            stacks.append(torch.from_numpy(ch).to(dev))
            flats.append(fd)
            where.append((i, c))
    for rep in range(2):
        got = gp.run_many(stacks, flats)
        for (i, c), g in zip(where, got):
            assert np.array_equal(g.cpu().numpy(), alone[i][c]), (rep, i, c)
    assert len(gp._graphs) == 2


def test_server_graph_replays_equal_alone(fovs, pipe, alone, dev):
    """Slots that replay their batches' kernel chains as hipGraphs (``HybServer(graph=True)``): two
    runs of the same requests (the second replays the graphs the first captured) give every
    channel's spots alone."""
    from mermake_amd.serve import FovRequest, HybServer
    data, _ = fovs
    srv = HybServer(pipe, batch_tiles=5, slots=2, graph=True)
    for rep in range(2):
        reqs = [FovRequest(i, [torch.from_numpy(c).pin_memory() for c in chans], torch.from_numpy(flat).pin_memory(),
                           arrival=0.003 * i) for i, (chans, flat) in enumerate(data)]
        done = srv.run(reqs)
        assert len(done) == len(reqs)
        for r in done:
            assert r.error is None
            for c, s in enumerate(r.spots):
                assert np.array_equal(s, alone[r.rid][c]), (rep, r.rid, c)
    graphs = sum(len(s.graphs) for s in srv.slots)
    assert 1 <= graphs <= 2 * 5
    assert srv.stats["graph_replays"] > 0


@pytest.mark.parametrize("host_inputs,batch_tiles,budget_fovs,cap_fovs",
                         [(False, 5, None, None), (True, 7, 1.5, None), (True, 36, None, None), (True, 7, None, 1.5),
                          (True, 7, None, 0.5)])
def test_server_mix_equals_alone(fovs, pipe, alone, dev, host_inputs, batch_tiles, budget_fovs, cap_fovs):
    """Requests arriving while others run are admitted as memory frees up (within the HBM budget and the
    input arena) and batched together; every request's spots equal its channels run alone.  An arena
    smaller than a request (cap 0.5) gives each request memory of its own, one request at a time."""
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
    per = max(sum(c.nbytes for c in chans) + flat.nbytes for chans, flat in data)
    budget = None if budget_fovs is None else int(budget_fovs * per)
    kw = {} if cap_fovs is None else {"inflight_cap": int(cap_fovs * per)}
    srv = HybServer(pipe, batch_tiles=batch_tiles, slots=2, hbm_budget=budget, **kw)
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
    if budget_fovs is not None or cap_fovs is not None:
        # with room for one request at a time, admissions follow completions
        order = sorted(done, key=lambda r: r.t_admit)
        for a, b in zip(order, order[1:]):
            assert b.t_admit >= a.t_done - 1e-9


def test_server_fails_bad_requests_alone(pipe, fovs, alone, dev):
    """Requests that cannot be served (over the HBM budget, wrong depth, a flat of the wrong shape, an
    unknown PSF mode, a flat list of the wrong length) fail alone, with the reason in ``error``; the
    others, arriving before, between and after them, give the spots of their channels alone."""
    from mermake_amd.serve import FovRequest, HybServer
    data, _ = fovs
    per = max(sum(c.nbytes for c in chans) + flat.nbytes for chans, flat in data)
    good = [FovRequest(i, [torch.from_numpy(c).pin_memory() for c in chans], torch.from_numpy(flat).pin_memory(),
                       arrival=0.002 * i) for i, (chans, flat) in enumerate(data)]
    chans0, flat0 = data[0]
    big = [torch.from_numpy(np.concatenate([c, c], axis=1)) for c in chans0 * 2]          # over the budget
    bad = [FovRequest(100, big, None, arrival=0.0),
           FovRequest(101, [torch.from_numpy(c[:5]) for c in chans0], torch.from_numpy(flat0), arrival=0.003),
           FovRequest(102, [torch.from_numpy(c) for c in chans0], torch.from_numpy(flat0[:, :7]), arrival=0.003),
           FovRequest(103, [torch.from_numpy(c) for c in chans0], torch.from_numpy(flat0), arrival=0.004,
                      psf="nope"),
           FovRequest(104, [torch.from_numpy(c) for c in chans0], [torch.from_numpy(flat0)], arrival=0.01)]
    srv = HybServer(pipe, batch_tiles=5, slots=2, hbm_budget=int(2.5 * per))
    done = srv.run(good + bad)
    assert len(done) == len(good) + len(bad)
    for r in bad:
        assert r.error and r.spots is None and r.t_done is not None, (r.rid, r.error)
    assert "budget" in bad[0].error and "depth" in bad[1].error and "flat" in bad[2].error
    assert "PSF mode" in bad[3].error and "flats for" in bad[4].error
    for r in good:
        assert r.error is None
        for c, sp in enumerate(r.spots):
            assert np.array_equal(sp, alone[r.rid][c]), (r.rid, c)
    assert srv.stats["failed"] == len(bad) and srv.stats["requests"] == len(good) + len(bad)
    assert srv.inflight_bytes == 0


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


def _dapi_pipe(depth, dev):
    from mermake_amd.deconvolver import Deconvolver
    from mermake_amd.pipeline import DapiPipeline
    from mermake_amd.synth import gaussian_psf
    psf = gaussian_psf((depth, 31, 31), 1.5, 1.2)
    d = Deconvolver(psf, (depth, 300, 300), tile_size=100, zpad=(depth - 1) // 2, overlap=20, beta=0.01)
    return DapiPipeline(d, blur_radius=50, threshold=1.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5)


def test_server_mixed_psf_modes_dapi_and_drift_equal_alone(fovs, pipe, alone, dev):
    """One server holds a single-PSF mode and a keyed multi-PSF mode, serves each request's DAPI channel and
    the drift of its FOV block: every request's hyb spots equal its mode's pipeline alone, its DAPI features
    equal DapiPipeline alone, and its drift equals ``DualAligner(block[len // 2], th)`` alone
    (``__main__.py:213-229,272-282``)."""
    import types
    from mermake_amd.align import DualAligner
    from mermake_amd.pipeline import HybPipeline
    from mermake_amd.serve import FovRequest, HybServer
    from mermake_amd.synth import gaussian_psf, nuclei_torch
    data, _ = fovs
    psfs = {(0, x, y): gaussian_psf((11, 31, 31), 1.2 + 0.2 * i, 1.0 + 0.15 * i)
            for i, (x, y) in enumerate(((75, 75), (75, 225), (225, 75), (225, 225)))}
    grid = HybPipeline.from_settings(psfs, SHAPES[0], tile_size=100, overlap=20, beta=1e-4, psf_mode="keyed", **HYB)
    dapi = _dapi_pipe(12, dev)
    modes = ["single", "grid", "grid", "single"]
    fov_of = ["A", "A", "B", "B"]
    shifts = [(0, 0, 0), (1, 4, -3), (0, 0, 0), (-1, -5, 6)]
    dapis = [nuclei_torch(50 + (0 if f == "A" else 1), (12,) + SHAPES[0 if f == "A" else 2][1:], 6, dev, shift=sh,
                          noise_seed=60 + i, radius=(12.0, 20.0), zradius=(3.0, 5.0), spots=20)
             for i, (f, sh) in enumerate(zip(fov_of, shifts))]
    want_spots, want_dapi = [], []
    for i, (chans, flat) in enumerate(data):
        p = pipe if modes[i] == "single" else grid
        fl = torch.from_numpy(flat).to(dev)
        want_spots.append([p(torch.from_numpy(c).to(dev), fl).cpu().numpy() for c in chans])
        pl, mi = dapi(dapis[i])
        want_dapi.append((pl.cpu().numpy(), mi.cpu().numpy()))
    reqs = [FovRequest(i, [torch.from_numpy(c).pin_memory() for c in chans], torch.from_numpy(flat).pin_memory(),
                       arrival=0.001 * i, dapi=dapis[i].cpu().pin_memory(), psf=modes[i], fov=fov_of[i],
                       last=(i == 3)) for i, (chans, flat) in enumerate(data)]
    print("DAPI features", [(len(p), len(m)) for p, m in want_dapi], flush=True)
    # a small scene: the drift threshold sits just above the DAPI threshold so that both signs keep points
    srv = HybServer({"single": pipe, "grid": grid}, batch_tiles={"single": 6, "grid": 5}, slots=2, dapi=dapi,
                    drift_th=1.2)
    done = srv.run(reqs)
    for r in done:
        assert r.error is None, r.error
        for c, sp in enumerate(r.spots):
            assert np.array_equal(sp, want_spots[r.rid][c]), (r.rid, c)
        assert np.array_equal(r.Xh_plus, want_dapi[r.rid][0]) and np.array_equal(r.Xh_minus, want_dapi[r.rid][1])
    st = srv.stats
    assert st["dapi_stacks"] == 4 and st["drift_blocks"] == 2
    assert st["tiles_per_mode"]["single"] > 0 and st["tiles_per_mode"]["grid"] > 0

    def feats(i):
        return types.SimpleNamespace(Xh_plus=torch.from_numpy(want_dapi[i][0]).to(dev),
                                     Xh_minus=torch.from_numpy(want_dapi[i][1]).to(dev))
    for members in ([0, 1], [2, 3]):
        ref = members[len(members) // 2]
        dual = DualAligner(feats(ref), th=1.2)
        for i in members:
            f, tp, tm, n_p, n_m = dual.get_best_translation_pointsV2(feats(i))
            g = done[i].drift
            assert done[i].drift_ref == ref
            assert np.array_equal(f, g[0]) and np.array_equal(tp, g[1]) and np.array_equal(tm, g[2])
            assert (n_p, n_m) == (g[3], g[4])
