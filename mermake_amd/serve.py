"""Served form of the FOV path: a continuous-batching request loop over FOV stacks.

A request is one FOV as the microscope delivers it (the per-image body of
``mermake/__main__.py:211-284``): hyb channel stacks (uint16, float32 or ``SquaredU8``; host or
device) with their flat fields, optionally the DAPI channel, a PSF mode and the FOV block it
belongs to.  Requests arrive over time.  The loop

* admits an arrived request when its host inputs fit the free room of the server's input arena
  (device memory allocated once at start, ``inflight_cap`` bytes within the HBM budget: free memory
  after the server's own buffers, minus a reserve), and copies them into it on a copy stream of
  its own.  A request that cannot
  be served (wrong rank or depth, flat of the wrong shape, unknown PSF mode, over the budget, a
  failure while its tiles are prepared) fails alone: its ``error`` is set, its bytes are released
  and the loop goes on, as the reference's image queue skips a file it cannot read
  (``io.py:451-460``);
* keeps one FIFO of tile jobs per PSF mode (each mode is a ``HybPipeline`` with its own resident
  Wiener spectra: one PSF, or a keyed PSF grid) over every admitted request, and fills each launch
  with up to ``batch_tiles`` jobs of the mode whose oldest job arrived first, so tiles of different
  requests and FOV sizes share launches;
* runs up to ``slots`` batches at once, each on its own stream with its own buffers: the
  deconvolution of one batch overlaps the maxima / fit of another and the host's readback;
* runs each request's DAPI channel through a ``DapiPipeline`` (``__main__.py:213-229``) on a stream
  of its own as soon as its copy has landed, and the drift (``align.drift``, ``__main__.py:272-282``)
  of a FOV block once the block is closed (a later request of another FOV, or ``last``) and its
  members are done.  Both run on a host thread of their own, so that their host synchronisations
  do not hold up the batch loop.  The block's ``DualAligner`` keeps the reference raster spectra resident for
  all of its members;
* reads results back asynchronously into pinned memory (counts first, then exactly the kept rows)
  and completes a request when its last tile and its DAPI features are back, recording the latency
  from arrival to the last spot array on the host.

Each channel's spots equal ``HybPipeline(stack, flat)`` of the request's PSF mode (the hyb loop of
``__main__.py:242-269``), its DAPI features equal ``DapiPipeline(dapi, flat)``, and its drift equals
``DualAligner(block[len(block) // 2], th=4).get_best_translation_pointsV2(request)``: a tile is
computed by the same kernels whatever batch it lands in, and rows are assembled in tile order.
"""
import collections
import concurrent.futures
import time
import types

import numpy as np
import torch

from . import _lib as L
from ._dev import capture_graph
from .deconvolver import prepare_flat, prepare_stack, tile_records

_REC = L.TILE_DTYPE.itemsize


class FovRequest:
    """One FOV.

    ``channels``: hyb channel stacks (z, x, y); ``flat``: one 2D flat for all of them, a list with one
    per channel (``None`` entries allowed), or ``None``; ``dapi`` / ``dapi_flat``: the DAPI channel and
    its flat (optional); ``psf``: key of the server's PSF mode for the hyb channels (``None``: the
    first); ``fov``: FOV block key for the drift (``None``: no drift); ``last``: closes the block;
    ``arrival``: seconds after the server's start.

    Filled by the server: ``spots`` (one (N, 8) float32 array per hyb channel), ``Xh_plus`` /
    ``Xh_minus`` (DAPI features), ``drift`` (``(tzxyf, t_plus, t_minus, N_plus, N_minus)``) and
    ``drift_ref`` (the rid of the block's reference), ``error`` (the reason a request failed, else
    ``None``), ``t_admit`` / ``t_done``."""

    def __init__(self, rid, channels, flat=None, arrival=0.0, dapi=None, dapi_flat=None, psf=None, fov=None,
                 last=False):
        self.rid = rid
        self.channels = list(channels)
        self.flat = flat
        self.dapi = dapi
        self.dapi_flat = dapi_flat
        self.psf = psf
        self.fov = fov
        self.last = bool(last)
        self.arrival = float(arrival)
        self.t_admit = None
        self.t_done = None
        self.spots = None
        self.Xh_plus = None
        self.Xh_minus = None
        self.drift = None
        self.drift_ref = None
        self.drift_error = None
        self.error = None
        self._mode = None
        self._chan = None     # device channels
        self._flats = None    # device flat of each channel (or None)
        self._dapi = None     # (device DAPI stack, device flat)
        self._evs = None      # copy-done event of each channel
        self._dapi_ev = None
        self._dapi_fut = None
        self._dapi_done = True
        self._next = 0        # channels whose tiles are queued
        self._rows = None
        self._left = 0
        self._bytes = 0

    def channel_flats(self):
        if isinstance(self.flat, (list, tuple)):
            if len(self.flat) != len(self.channels):
                raise ValueError(f"{len(self.flat)} flats for {len(self.channels)} channels")
            return list(self.flat)
        return [self.flat] * len(self.channels)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class _Job:
    __slots__ = ("req", "ch", "ti", "rec", "mode")

    def __init__(self, req, ch, ti, rec, mode):
        self.req, self.ch, self.ti, self.rec, self.mode = req, ch, ti, rec, mode


def _fused_blur(pipe):
    return pipe.blur_radius is not None and int(pipe.blur_radius) // 2 in (15, 25)


class _Arena:
    """Device memory for the inputs of the requests in flight, allocated (and so first touched) once
    when the server starts: each admitted request takes one block of it, in admission order, as a
    ring.  Blocks are freed in any order; the ring's tail passes a block once it and every block
    before it are free."""

    ALIGN = 256

    def __init__(self, nbytes, device):
        self.size = max(0, int(nbytes)) // self.ALIGN * self.ALIGN
        self.buf = torch.empty(self.size, dtype=torch.uint8, device=device) if self.size else None
        if self.buf is not None:
            self.buf.zero_()
        self.blocks = collections.deque()   # [start, end, free]
        self.head = 0

    @classmethod
    def rounded(cls, n):
        return -(-int(n) // cls.ALIGN) * cls.ALIGN

    def alloc(self, n):
        """A block of ``n`` bytes (a multiple of ALIGN) as [start, end, free], or None if it does not fit now."""
        if n > self.size:
            return None
        if not self.blocks:
            self.head = 0
            start = 0
        else:
            tail = self.blocks[0][0]
            if self.head > tail:                    # live blocks in [tail, head): room at the end, then the start
                if self.size - self.head >= n:
                    start = self.head
                elif n <= tail:
                    start = 0
                else:
                    return None
            elif tail - self.head >= n:             # wrapped: the room is [head, tail)
                start = self.head
            else:
                return None
        blk = [start, start + n, False]
        self.blocks.append(blk)
        self.head = start + n
        return blk

    def free(self, blk):
        blk[2] = True
        while self.blocks and self.blocks[0][2]:
            self.blocks.popleft()

    def view(self, blk, off, t):
        """Device tensor of ``t``'s dtype and shape at byte ``off`` of the block (keeps ``t``'s tensor type)."""
        n = t.numel() * t.element_size()
        v = self.buf[blk[0] + off:blk[0] + off + n].view(t.dtype).view(t.shape)
        return v if type(t) is torch.Tensor else v.as_subclass(type(t))


class _Slot:
    """Device and pinned buffers of one in-flight batch, sized for the largest batch of any mode."""

    def __init__(self, pipes, B_of, cap, device, graph=False):
        self.cap = cap
        self.B = max(B_of.values())
        self.device = device
        self.graphs = {} if graph else None
        self.replays = 0
        self.cap_stream = torch.cuda.Stream(device) if graph else None
        self.stream = torch.cuda.Stream(device)
        spec = max(B_of[k] * p.d._plan.spec_bytes for k, p in pipes.items())
        res = max(B_of[k] * p.d.tile_height * p.d._plan.T ** 2 for k, p in pipes.items())
        self.spec = torch.empty(spec, dtype=torch.uint8, device=device)
        self.res = torch.empty(res, dtype=torch.float32, device=device)
        self.tmp = None
        if any(p.blur_radius is not None and not p.spectral_blur() and not _fused_blur(p) for p in pipes.values()):
            self.tmp = torch.empty(res, dtype=torch.float32, device=device)
        self.mws_bytes = max(L.lib().mm_maxima_ws_bytes(p.d.tile_height, p.d._plan.T, p.d._plan.T, B_of[k])
                             for k, p in pipes.items())
        self.mws = torch.empty(self.mws_bytes, dtype=torch.uint8, device=device)
        B = self.B
        self.coords = torch.empty((B, cap, 3), dtype=torch.int32, device=device)
        self.fit = torch.empty((B, cap, 8), dtype=torch.float32, device=device)
        self.stats = torch.empty((2 * B + 1,), dtype=torch.int32, device=device)  # counts | kept | total
        self.xh = torch.empty((B * cap, 8), dtype=torch.float32, device=device)
        self.table = torch.empty((B * _REC,), dtype=torch.uint8, device=device)
        self.table_h = torch.empty((B * _REC,), dtype=torch.uint8, pin_memory=True)
        self.stats_h = torch.empty((2 * B + 1,), dtype=torch.int32, pin_memory=True)
        self.xh_h = torch.empty((B * cap, 8), dtype=torch.float32, pin_memory=True)
        self.ev_stats = torch.cuda.Event()
        self.ev_rows = torch.cuda.Event()
        self.jobs = None
        self.pipe = None
        self.state = "free"
        self.n = 0
        self.total = 0

    def _chain(self, pipe, n):
        """Enqueue deconvolution -> blur -> maxima -> fit -> keep of the ``n`` descriptors in the slot's table
        on the current stream, into the slot's buffers."""
        d = pipe.d
        T, sz = d._plan.T, d.tile_height
        single = d.psf_fft.shape[0] == 1
        st = L.stream_ptr(self.device)
        counts, kept, total = self.stats[:n], self.stats[self.B:self.B + n], self.stats[2 * self.B:]
        res = self.res[:n * sz * T * T]
        tmp = None if self.tmp is None else self.tmp[:n * sz * T * T]
        # the windows of the keep filter, as in HybPipeline: the deconvolution, the blur and the
        # candidates only where the kept spots can come from
        win = pipe.windows()
        delta = None if pipe.blur_radius is None else int(pipe.blur_radius) // 2
        tab = L.ptr(self.table)
        # the blur subtraction rides on the Wiener operator where the windows allow (HybPipeline)
        spectral = pipe.spectral_blur()
        ops = d.box_spectra(pipe.blur_radius) if spectral else d.psf_fft
        dargs = (d._plan.h, tab, n, L.ptr(ops), ops.shape[0], n if single else 1, L.ptr(self.spec), L.ptr(res))
        if win is None or delta is None:
            L.call("mm_deconv_tiles", *dargs, st)
        elif spectral:  # deconvolution minus blur, the threshold bits from its inverse y transform, maxima
            L.call("mm_tiles_deconv_maxima_window", *dargs, win[0][0], win[0][1], pipe.threshold, pipe.delta,
                   win[1][0], win[1][1], L.ptr(self.coords), self.cap, L.ptr(counts), L.ptr(self.mws),
                   self.mws_bytes, st)
        else:
            L.call("mm_deconv_tiles_window", *dargs, max(0, win[0][0] - delta), min(T, win[0][1] + delta), st)
        out = res
        fused = win is not None and (spectral or d.fused_blur_maxima(pipe.blur_radius))
        if delta is not None and not spectral:
            # blur subtraction into the spectrum buffer, which is free again (deconvolver.py:160-163)
            out = self.spec[:n * sz * T * T * 4].view(torch.float32)
            if win is None:
                L.call("mm_blur_sub_xy", L.ptr(res), L.ptr(out), L.ptr(tmp), n * sz, T, delta, st)
            elif fused:  # blur + maxima threshold bits, then the maxima of the candidates
                L.call("mm_tiles_blur_maxima_window", d._plan.h, tab, n, L.ptr(res), L.ptr(out), delta,
                       win[0][0], win[0][1], pipe.threshold, pipe.delta, win[1][0], win[1][1],
                       L.ptr(self.coords), self.cap, L.ptr(counts), L.ptr(self.mws), self.mws_bytes, st)
            else:
                L.call("mm_blur_sub_xy_window", L.ptr(res), L.ptr(out), L.ptr(tmp), n * sz, T, delta,
                       win[0][0], win[0][1], st)
        if win is None:
            L.call("mm_tiles_maxima", d._plan.h, tab, n, L.ptr(out), pipe.threshold, pipe.delta,
                   L.ptr(self.coords), self.cap, L.ptr(counts), L.ptr(self.mws), self.mws_bytes, st)
        elif not fused:
            L.call("mm_tiles_maxima_window", d._plan.h, tab, n, L.ptr(out), pipe.threshold, pipe.delta,
                   win[1][0], win[1][1], L.ptr(self.coords), self.cap, L.ptr(counts), L.ptr(self.mws),
                   self.mws_bytes, st)
        L.call("mm_tiles_fit", d._plan.h, tab, n, L.ptr(out), L.ptr(self.coords), L.ptr(counts), self.cap,
               pipe.delta_fit, pipe.sigmaZ, pipe.sigmaXY, L.ptr(self.fit), st)
        L.call("mm_tiles_keep", tab, n, d.overlap, L.ptr(self.fit), L.ptr(counts), self.cap, L.ptr(kept),
               L.ptr(self.xh), n * self.cap, L.ptr(total), None, st)

    def launch(self, pipe, jobs, capture=True):
        d = pipe.d
        n = len(jobs)
        T, sz = d._plan.T, d.tile_height
        if n > self.B or n * d._plan.spec_bytes > self.spec.numel() or n * sz * T * T > self.res.numel():
            raise RuntimeError("the slot's buffers do not fit this batch (was the schedule changed after the "
                               "server was made?)")
        recs = np.empty(n, dtype=L.TILE_DTYPE)
        for i, j in enumerate(jobs):
            recs[i] = j.rec
        self.table_h[:n * _REC].copy_(torch.from_numpy(recs.view(np.uint8)))
        self.jobs, self.n, self.pipe, self.state = jobs, n, pipe, "stats"
        # one hipGraph per (pipeline, schedule, tiles): the first batch of a shape runs eagerly, the graph
        # is captured after it and replayed by the next ones (the table is the graph's input); the server
        # captures full batches only, so that partial batches of every size do not each add a graph
        key = (id(pipe), d.generation, n)
        g = None if self.graphs is None else self.graphs.get(key)
        with torch.cuda.stream(self.stream):
            self.table[:n * _REC].copy_(self.table_h[:n * _REC], non_blocking=True)
            if g is not None:
                g.replay()
                self.replays += 1
            else:
                self._chain(pipe, n)
            self.stats_h.copy_(self.stats, non_blocking=True)
            self.ev_stats.record(self.stream)
        if self.graphs is not None and g is None and capture:
            self.graphs[key] = capture_graph(lambda: self._chain(pipe, n), self.cap_stream)

    def fetch_rows(self):
        """After the counts are back: copy exactly the kept rows (or report an overflowing tile)."""
        s = self.stats_h.numpy()
        if int(s[:self.n].max(initial=0)) > self.cap:
            return False
        self.total = int(s[2 * self.B])
        with torch.cuda.stream(self.stream):
            if self.total:
                self.xh_h[:self.total].copy_(self.xh[:self.total], non_blocking=True)
            self.ev_rows.record(self.stream)
        self.state = "rows"
        return True


def _pctl(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(np.ceil(q * len(v))) - 1)] if v else None


class HybServer:
    """Continuous-batching request loop over one or several PSF modes (``HybPipeline`` each), with an
    optional DAPI pipeline and the drift of FOV blocks.

    ``pipe``: a ``HybPipeline``, or a dict {mode key: HybPipeline} (requests choose with ``psf=key``);
    ``batch_tiles``: tiles per launch, an int or a dict per mode; ``dapi``: a ``DapiPipeline`` for the
    requests' DAPI channels; ``drift_th``: the threshold of ``align.drift`` (``align.py:169``); ``graph``:
    each slot replays its full batches' kernel chain as a hipGraph, one per mode;
    ``inflight_cap``: bytes of the device arena that holds the host inputs of the requests in flight
    (allocated once, here, within the HBM budget).  Inputs admitted far ahead of the GPU only wait; and memory allocated fresh
    while the loop runs is cleared by the driver on the DMA engines that also copy the inputs, which
    halves their rate on a fresh GPU.  A request larger than the arena gets memory of its own within the
    HBM budget and runs when no other such request is in flight."""

    def __init__(self, pipe, batch_tiles=36, slots=2, cap=None, hbm_reserve=4 << 30, hbm_budget=None, dapi=None,
                 drift_th=4, graph=False, inflight_cap=16 << 30):
        self.pipes = dict(pipe) if isinstance(pipe, dict) else {None: pipe}
        if not self.pipes:
            raise ValueError("HybServer needs at least one pipeline")
        for p in self.pipes.values():
            p.prepare()     # box-folded operators before the first request, not inside its batch
        if dapi is not None:
            dapi.prepare()
        self.default_mode = next(iter(self.pipes))
        self.pipe = self.pipes[self.default_mode]
        devs = {p.d.device for p in self.pipes.values()}
        if len(devs) != 1:
            raise ValueError("every pipeline of a server must live on one device")
        self.device = devs.pop()
        self.B_of = {k: int(batch_tiles[k] if isinstance(batch_tiles, dict) else batch_tiles) for k in self.pipes}
        self.B = max(self.B_of.values())
        self.cap = int(cap or max(p.cap for p in self.pipes.values()))
        self.slots = [_Slot(self.pipes, self.B_of, self.cap, self.device, graph) for _ in range(slots)]
        self.copy_stream = torch.cuda.Stream(self.device)
        self.dapi = dapi
        self.dapi_stream = torch.cuda.Stream(self.device) if dapi is not None else None
        self.drift_th = drift_th
        torch.cuda.synchronize(self.device)
        free, _ = torch.cuda.mem_get_info(self.device)
        self.budget = int(hbm_budget) if hbm_budget is not None else max(0, free - int(hbm_reserve))
        self.arena = _Arena(min(int(inflight_cap), self.budget), self.device)
        self.inflight_bytes = 0      # bytes of a request in flight outside the arena (larger than it)
        self.stats = {}

    # ------------------------------------------------------------------ admission
    @staticmethod
    def _host_bytes(t):
        if t is None:
            return 0
        if isinstance(t, torch.Tensor):
            return 0 if t.is_cuda else t.numel() * t.element_size()
        return np.asarray(t).nbytes

    def _host_inputs(self, req):
        """The request's host-resident inputs, each once (a flat shared by several channels is copied once)."""
        seen, out = set(), []
        for t in list(req.channels) + list(req.channel_flats()) + [req.dapi, req.dapi_flat]:
            if t is not None and id(t) not in seen and self._host_bytes(t):
                seen.add(id(t))
                out.append(t)
        return out

    def _need(self, req):
        """Device bytes of the request's host inputs (each rounded to the arena's alignment)."""
        return sum(_Arena.rounded(self._host_bytes(t)) for t in self._host_inputs(req))

    @staticmethod
    def _shape(t):
        return tuple(t.shape) if hasattr(t, "shape") else np.asarray(t).shape

    def _validate(self, req):
        """The checks a request must pass before any of its work is queued (ValueError otherwise)."""
        if req.psf not in self.pipes and not (req.psf is None and self.default_mode in self.pipes):
            raise ValueError(f"unknown PSF mode {req.psf!r} (the server has {list(self.pipes)})")
        d = self.pipes[req.psf if req.psf in self.pipes else self.default_mode].d
        for c, f in zip(req.channels, req.channel_flats()):
            s = self._shape(c)
            if len(s) != 3 or s[0] != d.tile_height:
                raise ValueError(f"channel of shape {s}: the pipeline takes (z, x, y) stacks of depth {d.tile_height}")
            if f is not None:
                fs = self._shape(f)
                if fs not in (s[1:], (1,) + s[1:]):
                    raise ValueError(f"flat field shape {fs} does not match the stack's {s[1:]}")
        if req.dapi is not None:
            if self.dapi is None:
                raise ValueError("a DAPI channel was sent to a server without a DAPI pipeline")
            s = self._shape(req.dapi)
            if len(s) != 3 or s[0] != self.dapi.d.tile_height:
                raise ValueError(f"DAPI stack of shape {s}: the DAPI pipeline takes depth {self.dapi.d.tile_height}")
            if req.dapi_flat is not None and self._shape(req.dapi_flat) not in (s[1:], (1,) + s[1:]):
                raise ValueError(f"DAPI flat shape {self._shape(req.dapi_flat)} does not match {s[1:]}")
        need = self._need(req)
        if need > self.budget:
            raise ValueError(f"request {req.rid} needs {need} bytes of HBM, over the budget {self.budget}")

    def _to_device(self, t, req=None):
        """``t`` on the device: device tensors as they are; host inputs copied (on the current stream) into
        the request's arena block, or into memory of their own for a request larger than the arena."""
        if t is None:
            return None
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t
        h = t if isinstance(t, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(t))
        if req is None or req._blk is None:
            return h.to(self.device, non_blocking=True)
        if not h.is_contiguous():
            h = h.contiguous()
        off = req._placed.get(id(t))
        if off is None:
            off = req._fill
            req._placed[id(t)] = off
            req._fill += _Arena.rounded(h.numel() * h.element_size())
        d = self.arena.view(req._blk, off, h)
        d.copy_(h, non_blocking=True)
        return d

    def _reserve(self, req):
        """Device memory for the request's host inputs: a block of the arena, or, for a request larger
        than the arena (and within the HBM budget, which ``_validate`` checks), memory of its own while no
        other request is in flight.  False when it has to wait."""
        need = self._need(req)
        req._blk, req._bytes = None, 0
        if need == 0:
            return True
        blk = self.arena.alloc(need)
        if blk is not None:
            req._blk = blk
            return True
        if need <= self.arena.size or self.inflight_bytes or self.arena.blocks:
            return False                 # room comes back as requests in flight complete
        req._bytes = need
        self.inflight_bytes += need
        return True

    def _admit(self, req, now):
        """Start the request's host-to-device copies on the copy stream, each channel's flat before the
        channel, with an event after each channel: a channel's tiles are queued as soon as its own copy is
        done."""
        req.t_admit = now
        req._rows = [None] * len(req.channels)
        req._left = 0
        req._next = 0
        req._dapi_done = req.dapi is None
        req._mode = req.psf if req.psf in self.pipes else self.default_mode
        req._placed, req._fill = {}, 0
        copied = {}
        chans, flats, evs = [], [], []
        with torch.cuda.stream(self.copy_stream):
            for c, f in zip(req.channels, req.channel_flats()):
                if f is not None and id(f) not in copied:
                    copied[id(f)] = self._to_device(f, req)
                flats.append(None if f is None else copied[id(f)])
                chans.append(self._to_device(c, req))
                ev = torch.cuda.Event()
                ev.record(self.copy_stream)
                evs.append(ev)
            if req.dapi is not None:
                req._dapi = (self._to_device(req.dapi, req), self._to_device(req.dapi_flat, req))
                req._dapi_ev = torch.cuda.Event()
                req._dapi_ev.record(self.copy_stream)
        req._chan, req._flats, req._evs = chans, flats, evs

    def _enqueue_channel(self, req, ch, fifo):
        d = self.pipes[req._mode].d
        stack, flat = req._chan[ch], req._flats[ch]
        d._check(stack)
        img = prepare_stack(stack)
        f = prepare_flat(flat, img)
        req._chan[ch] = img
        req._flats[ch] = f           # keeps a cast or reshaped flat alive with the request
        recs = tile_records(d.geometry(img.shape), img, f)
        req._rows[ch] = [None] * len(recs)
        req._left += len(recs)
        for ti in range(len(recs)):
            fifo.append(_Job(req, ch, ti, recs[ti], req._mode))

    def _dapi_work(self, stack, flat, ev):
        """The DAPI block of ``__main__.py:217-225`` for one request, on the DAPI stream (runs on the
        server's DAPI thread: its host synchronisations do not hold up the batch loop)."""
        with torch.cuda.device(self.device), torch.cuda.stream(self.dapi_stream):
            self.dapi_stream.wait_event(ev)
            plus, minus = self.dapi(stack, flat)
            return plus.cpu().numpy(), minus.cpu().numpy()

    def _dapi_collect(self, req):
        """The DAPI features once the DAPI thread has them (its exception fails the request)."""
        plus, minus = req._dapi_fut.result()
        req.Xh_plus, req.Xh_minus = plus, minus
        req._dapi = None
        req._dapi_fut = None
        req._dapi_done = True

    # ------------------------------------------------------------------ completion
    def _release(self, req):
        if getattr(req, "_blk", None) is not None:
            self.arena.free(req._blk)
            req._blk = None
        if getattr(req, "_bytes", 0):
            self.inflight_bytes -= req._bytes
            req._bytes = 0
        req._chan = req._flats = req._dapi = req._evs = req._dapi_ev = None

    def _fail(self, req, err, now, done):
        if req.t_done is not None:
            return
        req.error = f"{type(err).__name__}: {err}" if isinstance(err, BaseException) else str(err)
        req.spots = None
        req._rows = None
        req._left = 0
        req._dapi_done = True
        req._dapi_fut = None
        req.t_done = now
        self._release(req)
        done.append(req)

    def _complete(self, req, now, done):
        """A request is done when every channel's tiles are queued and finished and its DAPI features
        are back."""
        if req.t_done is not None or req._left or req._next < len(req.channels) or not req._dapi_done:
            return
        req.spots = [np.concatenate(r) if r else np.zeros((0, 8), np.float32) for r in req._rows]
        req.t_done = now
        req._rows = None
        self._release(req)
        done.append(req)

    def _finish_tile(self, job, rows, now, done):
        req = job.req
        if req.error is not None:
            return
        req._rows[job.ch][job.ti] = rows
        req._left -= 1
        self._complete(req, now, done)

    def _redo_alone(self, jobs, now, done):
        """Each job alone through its pipeline (a batch overflowed the candidate cap, or failed); a job that
        fails again fails its request."""
        for job in jobs:
            if job.req.error is not None:
                continue
            recs = np.empty(1, dtype=L.TILE_DTYPE)
            recs[0] = job.rec
            try:
                table = torch.from_numpy(recs.view(np.uint8).copy()).to(self.device)
                rows = self.pipes[job.mode].tile_rows(table, 1)[0]
            except Exception as e:  # noqa: BLE001 - a request fails alone
                self._fail(job.req, e, now, done)
                continue
            self._finish_tile(job, rows, now, done)

    def _parse(self, slot, now, done):
        s = slot.stats_h.numpy()
        kept = s[slot.B:slot.B + slot.n]
        rows = slot.xh_h[:slot.total].numpy()
        off = 0
        for i, job in enumerate(slot.jobs):
            k = int(kept[i])
            self._finish_tile(job, rows[off:off + k].copy(), now, done)
            off += k

    # ------------------------------------------------------------------ drift of FOV blocks
    def _run_drift(self, block):
        """``align.drift`` (``align.py:163-177``) over a closed block's served members: the reference is
        the middle member, and its DualAligner (resident raster spectra) serves every member."""
        from .align import DualAligner
        members = [r for r in block if r.error is None and r.Xh_plus is not None]
        if not members:
            return
        ref = members[len(members) // 2]
        t0 = time.perf_counter()
        with torch.cuda.device(self.device), torch.cuda.stream(self.dapi_stream):
            def feats(r):
                return types.SimpleNamespace(Xh_plus=torch.from_numpy(r.Xh_plus).to(self.device),
                                             Xh_minus=torch.from_numpy(r.Xh_minus).to(self.device))
            try:
                dual = DualAligner(feats(ref), th=self.drift_th)
                for r in members:
                    r.drift = dual.get_best_translation_pointsV2(feats(r))
                    r.drift_ref = ref.rid
            except Exception as e:  # noqa: BLE001 - the block's drift fails, its members' spots stand
                for r in members:
                    r.drift = None
                    r.drift_error = f"{type(e).__name__}: {e}"
        self._drift_s.append(time.perf_counter() - t0)

    # ------------------------------------------------------------------ loop
    def run(self, requests):
        """Serve ``requests`` (arrival times relative to now); returns them completed, with ``spots``,
        ``Xh_plus`` / ``Xh_minus``, ``drift``, ``error``, ``t_admit`` and ``t_done`` set, and fills
        ``self.stats``."""
        aux = concurrent.futures.ThreadPoolExecutor(max_workers=1, thread_name_prefix="mm-dapi")
        try:
            return self._run(requests, aux)
        finally:
            aux.shutdown(wait=True)

    def _run(self, requests, aux):
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        nxt = 0
        waiting = collections.deque()
        copying = []
        fifos = {k: collections.deque() for k in self.pipes}
        done = []
        blocks = []          # [fov, members, closed]
        self._drift_s = []
        launches = 0
        tiles = 0
        overflows = 0
        tiles_of = {k: 0 for k in self.pipes}

        def close_blocks(keep_open=None):
            for b in blocks:
                if b[0] != keep_open and b[2] is False:
                    b[2] = True

        while True:
            now = time.perf_counter() - t0
            progressed = False
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                r = reqs[nxt]
                nxt += 1
                if r.fov is not None:
                    # a request of another FOV closes the open block (Block.add, io.py:302-310)
                    close_blocks(keep_open=r.fov)
                    if not blocks or blocks[-1][0] != r.fov or blocks[-1][2]:
                        blocks.append([r.fov, [], False])
                    blocks[-1][1].append(r)
                    if r.last:
                        blocks[-1][2] = True
                try:
                    self._validate(r)
                except Exception as e:  # noqa: BLE001 - a request fails alone
                    self._fail(r, e, now, done)
                    progressed = True
                    continue
                waiting.append(r)
            while waiting and self._reserve(waiting[0]):
                r = waiting.popleft()
                try:
                    self._admit(r, now)
                except Exception as e:  # noqa: BLE001
                    self._fail(r, e, now, done)
                    continue
                copying.append(r)
                progressed = True
            for r in list(copying):
                if r.t_done is not None:
                    copying.remove(r)
                    continue
                try:
                    while r._next < len(r.channels) and r._evs[r._next].query():
                        self._enqueue_channel(r, r._next, fifos[r._mode])
                        r._next += 1
                        progressed = True
                    if not r._dapi_done and r._dapi_fut is None and r._dapi_ev.query():
                        r._dapi_fut = aux.submit(self._dapi_work, r._dapi[0], r._dapi[1], r._dapi_ev)
                        progressed = True
                    if r._dapi_fut is not None and r._dapi_fut.done():
                        self._dapi_collect(r)
                        progressed = True
                except Exception as e:  # noqa: BLE001 - a request fails alone
                    self._fail(r, e, time.perf_counter() - t0, done)
                    copying.remove(r)
                    progressed = True
                    continue
                if r._next == len(r.channels) and r._dapi_done:
                    copying.remove(r)
                    self._complete(r, time.perf_counter() - t0, done)  # requests without tiles
                    progressed = True
            for slot in self.slots:
                if slot.state == "free":
                    for q in fifos.values():   # jobs of failed requests are dropped
                        while q and q[0].req.error is not None:
                            q.popleft()
                    live = [k for k, q in fifos.items() if q]
                    if not live:
                        continue
                    mode = min(live, key=lambda k: (fifos[k][0].req.arrival, fifos[k][0].req.t_admit))
                    q = fifos[mode]
                    batch = []
                    while q and len(batch) < self.B_of[mode]:
                        j = q.popleft()
                        if j.req.error is None:
                            batch.append(j)
                    if not batch:
                        continue
                    try:
                        slot.launch(self.pipes[mode], batch, capture=len(batch) == self.B_of[mode])
                    except Exception:  # noqa: BLE001 - retried tile by tile; failing tiles fail their request
                        torch.cuda.synchronize(self.device)
                        slot.state = "free"
                        self._redo_alone(batch, time.perf_counter() - t0, done)
                        progressed = True
                        continue
                    launches += 1
                    tiles += len(batch)
                    tiles_of[mode] += len(batch)
                    progressed = True
                elif slot.state == "stats" and slot.ev_stats.query():
                    if not slot.fetch_rows():
                        overflows += 1
                        self._redo_alone(slot.jobs, time.perf_counter() - t0, done)
                        slot.state = "free"
                    progressed = True
                elif slot.state == "rows" and slot.ev_rows.query():
                    self._parse(slot, time.perf_counter() - t0, done)
                    slot.state = "free"
                    progressed = True
            if nxt == len(reqs) and not waiting:
                close_blocks()
            for b in blocks:
                if b[2] is True and all(r.t_done is not None for r in b[1]):
                    b[2] = aux.submit(self._run_drift, b[1])   # closed and done: drift on the DAPI thread
                    progressed = True
                elif b[2] not in (True, False) and b[2].done():
                    b[2].result()
                    b[2] = None          # drifted
                    progressed = True
            blocks = [b for b in blocks if b[2] is not None]
            if len(done) == len(reqs) and not blocks:
                break
            if not progressed:
                idle = not any(fifos.values()) and not copying and all(s.state == "free" for s in self.slots)
                if idle and not waiting and nxt < len(reqs):
                    time.sleep(max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0)))
                else:
                    time.sleep(20e-6)
        wall = time.perf_counter() - t0
        ok = [r for r in reqs if r.error is None]
        lat = [r.latency for r in ok]
        self.stats = {
            "requests": len(reqs), "failed": len(reqs) - len(ok), "wall_s": wall,
            "fovs_per_s": len(ok) / wall if wall > 0 else None,
            "p50_latency_s": _pctl(lat, 0.5), "p99_latency_s": _pctl(lat, 0.99),
            "launches": launches, "tiles": tiles, "tiles_per_mode": {str(k): v for k, v in tiles_of.items()},
            "overflow_batches": overflows, "batch_tiles": self.B, "slots": len(self.slots),
            "dapi_stacks": sum(1 for r in ok if r.Xh_plus is not None),
            "drift_blocks": len(self._drift_s), "drift_s_total": float(sum(self._drift_s)),
            "hbm_budget_bytes": self.budget, "arena_bytes": self.arena.size,
            "graph_replays": sum(s.replays for s in self.slots),
        }
        return reqs


def poisson_arrivals(n, rate, seed=0):
    """``n`` arrival times (s) of a Poisson process of ``rate`` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate > 0 else np.zeros(n)
