"""Served form of the hyb path: a continuous-batching request loop over FOV stacks.

A request is one FOV: several channel stacks (uint16 or float32, host or device) that share
a flat field.  Requests arrive over time, as from the microscope.  The loop

* admits an arrived request when its inputs fit the HBM budget (device memory free when the
  server starts, minus a reserve, minus the inputs of the requests still in flight), and
  copies host inputs to the device on a copy stream of its own;
* keeps one FIFO of tile jobs over every admitted request and fills each launch with up to
  ``batch_tiles`` of them, so tiles of different requests and FOV sizes share launches;
* runs up to ``slots`` batches at once, each on its own stream with its own buffers: the
  deconvolution of one batch overlaps the maxima / fit of another and the host's readback;
* reads results back asynchronously into pinned memory (counts first, then exactly the kept
  rows) and completes a request when its last tile is back, recording the latency from
  arrival to the last spot array on the host.

Each channel's spots equal ``HybPipeline(stack, flat)`` (the hyb loop of
``mermake/__main__.py:242-269``): a tile is computed by the same kernels whatever batch it
lands in, and rows are assembled in tile order.
"""
import collections
import ctypes
import time

import numpy as np
import torch

from . import _lib as L
from .deconvolver import prepare_flat, prepare_stack, tile_records

_REC = L.TILE_DTYPE.itemsize


class FovRequest:
    """One FOV: ``channels`` (each a (z, x, y) stack), an optional 2D ``flat`` and the
    ``arrival`` time in seconds after the server's start.  ``spots`` (one (N, 8) float32
    array per channel) and the timestamps are filled by the server."""

    def __init__(self, rid, channels, flat=None, arrival=0.0):
        self.rid = rid
        self.channels = list(channels)
        self.flat = flat
        self.arrival = float(arrival)
        self.t_admit = None
        self.t_done = None
        self.spots = None
        self._dev = None
        self._evs = None   # copy-done event of each channel (the flat is copied before channel 0)
        self._next = 0     # channels whose tiles are queued
        self._rows = None
        self._left = 0
        self._bytes = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class _Job:
    __slots__ = ("req", "ch", "ti", "rec")

    def __init__(self, req, ch, ti, rec):
        self.req, self.ch, self.ti, self.rec = req, ch, ti, rec


class _Slot:
    """Device and pinned buffers of one in-flight batch of up to ``B`` tiles."""

    def __init__(self, pipe, B, cap, device):
        d = pipe.d
        T, sz = d._plan.T, d.tile_height
        self.B, self.cap = B, cap
        self.stream = torch.cuda.Stream(device)
        self.spec = torch.empty(B * d._plan.spec_bytes, dtype=torch.uint8, device=device)
        self.res = torch.empty((B, sz, T, T), dtype=torch.float32, device=device)
        self.tmp = None
        if pipe.blur_radius is not None and int(pipe.blur_radius) // 2 not in (15, 25):
            self.tmp = torch.empty((B, sz, T, T), dtype=torch.float32, device=device)
        self.mws_bytes = L.lib().mm_maxima_ws_bytes(sz, T, T, B)
        self.mws = torch.empty(self.mws_bytes, dtype=torch.uint8, device=device)
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
        self.state = "free"
        self.n = 0
        self.total = 0

    def launch(self, pipe, jobs, single):
        d = pipe.d
        n = len(jobs)
        T, sz = d._plan.T, d.tile_height
        recs = np.empty(n, dtype=L.TILE_DTYPE)
        for i, j in enumerate(jobs):
            recs[i] = j.rec
        self.table_h[:n * _REC].copy_(torch.from_numpy(recs.view(np.uint8)))
        st = ctypes.c_void_p(self.stream.cuda_stream)
        counts, kept, total = self.stats[:n], self.stats[self.B:self.B + n], self.stats[2 * self.B:]
        # the windows of the keep filter, as in HybPipeline: the deconvolution, the blur and the
        # candidates only where the kept spots can come from
        win = pipe.windows()
        delta = None if pipe.blur_radius is None else int(pipe.blur_radius) // 2
        with torch.cuda.stream(self.stream):
            self.table[:n * _REC].copy_(self.table_h[:n * _REC], non_blocking=True)
            tab = L.ptr(self.table)
            dargs = (d._plan.h, tab, n, L.ptr(d.psf_fft), d.psf_fft.shape[0], n if single else 1, L.ptr(self.spec),
                     L.ptr(self.res))
            if win is None or delta is None:
                L.call("mm_deconv_tiles", *dargs, st)
            else:
                L.call("mm_deconv_tiles_window", *dargs, max(0, win[0][0] - delta), min(T, win[0][1] + delta), st)
            out = self.res
            fused = win is not None and d.fused_blur_maxima(pipe.blur_radius)
            if delta is not None:
                # blur subtraction into the spectrum buffer, which is free again (deconvolver.py:160-163)
                out = self.spec[:n * sz * T * T * 4].view(torch.float32)
                if win is None:
                    L.call("mm_blur_sub_xy", L.ptr(self.res), L.ptr(out), L.ptr(self.tmp), n * sz, T, delta, st)
                elif fused:  # blur + maxima threshold bits, then the maxima of the candidates
                    L.call("mm_tiles_blur_maxima_window", d._plan.h, tab, n, L.ptr(self.res), L.ptr(out), delta,
                           win[0][0], win[0][1], pipe.threshold, pipe.delta, win[1][0], win[1][1],
                           L.ptr(self.coords), self.cap, L.ptr(counts), L.ptr(self.mws), self.mws_bytes, st)
                else:
                    L.call("mm_blur_sub_xy_window", L.ptr(self.res), L.ptr(out), L.ptr(self.tmp), n * sz, T, delta,
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
            self.stats_h.copy_(self.stats, non_blocking=True)
            self.ev_stats.record(self.stream)
        self.jobs, self.n, self.state = jobs, n, "stats"

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


class HybServer:
    """Continuous-batching request loop over a :class:`~mermake_amd.pipeline.HybPipeline`."""

    def __init__(self, pipe, batch_tiles=36, slots=2, cap=None, hbm_reserve=4 << 30, hbm_budget=None):
        self.pipe = pipe
        d = pipe.d
        self.device = d.device
        self.B = int(batch_tiles)
        self.cap = int(cap or pipe.cap)
        self.slots = [_Slot(pipe, self.B, self.cap, self.device) for _ in range(slots)]
        self.copy_stream = torch.cuda.Stream(self.device)
        torch.cuda.synchronize(self.device)
        free, _ = torch.cuda.mem_get_info(self.device)
        self.budget = int(hbm_budget) if hbm_budget is not None else max(0, free - int(hbm_reserve))
        self.single = d.psf_fft.shape[0] == 1
        self.inflight_bytes = 0
        self.stats = {}

    # ------------------------------------------------------------------ admission
    @staticmethod
    def _host_bytes(t):
        if isinstance(t, torch.Tensor):
            return 0 if t.is_cuda else t.numel() * t.element_size()
        return np.asarray(t).nbytes

    def _need(self, req):
        n = sum(self._host_bytes(c) for c in req.channels)
        if req.flat is not None:
            n += self._host_bytes(req.flat)
        return n

    def _admit(self, req, now):
        """Start the request's host-to-device copies on the copy stream, the flat first, with an
        event after each channel: a channel's tiles are queued as soon as its own copy is done."""
        req._bytes = self._need(req)
        self.inflight_bytes += req._bytes
        req.t_admit = now
        d = self.pipe.d
        req._rows = [[None] * len(d.geometry(tuple(c.shape))) for c in req.channels]
        req._left = sum(len(r) for r in req._rows)
        req._next = 0
        dev, evs = [], []
        with torch.cuda.stream(self.copy_stream):
            for c in ([req.flat] if req.flat is not None else []) + req.channels:
                if isinstance(c, torch.Tensor) and c.is_cuda:
                    dev.append(c)
                else:
                    h = c if isinstance(c, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(c))
                    dev.append(h.to(self.device, non_blocking=True))
                if c is not req.flat:
                    ev = torch.cuda.Event()
                    ev.record(self.copy_stream)
                    evs.append(ev)
        req._evs = evs
        req._dev = dev  # [flat], channels...

    def _enqueue_channel(self, req, ch, jobs):
        d = self.pipe.d
        k = 1 if req.flat is not None else 0
        flat = req._dev[0] if k else None
        stack = req._dev[k + ch]
        d._check(stack)
        img = prepare_stack(stack)
        f = prepare_flat(flat, img)
        req._dev[k + ch] = img
        if f is not None and f is not flat:
            req._dev.append(f)
        recs = tile_records(d.geometry(img.shape), img, f)
        for ti in range(len(recs)):
            jobs.append(_Job(req, ch, ti, recs[ti]))

    # ------------------------------------------------------------------ completion
    def _complete(self, req, now, done):
        """A request is done when every channel's tiles are queued and finished."""
        if req.t_done is not None or req._left or req._next < len(req.channels):
            return
        req.spots = [np.concatenate(r) if r else np.zeros((0, 8), np.float32) for r in req._rows]
        req.t_done = now
        req._rows = None
        req._dev = None
        self.inflight_bytes -= req._bytes
        done.append(req)

    def _finish_tile(self, job, rows, now, done):
        req = job.req
        req._rows[job.ch][job.ti] = rows
        req._left -= 1
        self._complete(req, now, done)

    def _overflow(self, slot, now, done):
        """A tile had more candidates than the slot's cap: redo the batch's tiles one stack at a time."""
        for job in slot.jobs:
            recs = np.empty(1, dtype=L.TILE_DTYPE)
            recs[0] = job.rec
            table = torch.from_numpy(recs.view(np.uint8).copy()).to(self.device)
            rows = self.pipe.tile_rows(table, 1)[0]
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

    # ------------------------------------------------------------------ loop
    def run(self, requests):
        """Serve ``requests`` (arrival times relative to now); returns them completed, with
        ``spots``, ``t_admit`` and ``t_done`` set, and fills ``self.stats``."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if self._need(r) > self.budget:
                raise ValueError(f"request {r.rid} needs {self._need(r)} bytes of HBM, over the budget {self.budget}")
        t0 = time.perf_counter()
        nxt = 0
        waiting = collections.deque()
        copying = []
        jobs = collections.deque()
        done = []
        launches = 0
        tiles = 0
        overflows = 0
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            progressed = False
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            while waiting and self.inflight_bytes + self._need(waiting[0]) <= self.budget:
                r = waiting.popleft()
                self._admit(r, now)
                copying.append(r)
                progressed = True
            for r in list(copying):
                while r._next < len(r.channels) and r._evs[r._next].query():
                    self._enqueue_channel(r, r._next, jobs)
                    r._next += 1
                    progressed = True
                if r._next == len(r.channels):
                    copying.remove(r)
                    self._complete(r, time.perf_counter() - t0, done)  # requests without tiles
                    progressed = True
            for slot in self.slots:
                if slot.state == "free" and jobs:
                    batch = [jobs.popleft() for _ in range(min(self.B, len(jobs)))]
                    slot.launch(self.pipe, batch, self.single)
                    launches += 1
                    tiles += len(batch)
                    progressed = True
                elif slot.state == "stats" and slot.ev_stats.query():
                    if not slot.fetch_rows():
                        overflows += 1
                        self._overflow(slot, time.perf_counter() - t0, done)
                        slot.state = "free"
                    progressed = True
                elif slot.state == "rows" and slot.ev_rows.query():
                    self._parse(slot, time.perf_counter() - t0, done)
                    slot.state = "free"
                    progressed = True
            if not progressed:
                idle = not jobs and not copying and all(s.state == "free" for s in self.slots)
                if idle and not waiting and nxt < len(reqs):
                    time.sleep(max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0)))
                else:
                    time.sleep(20e-6)
        wall = time.perf_counter() - t0
        lat = sorted(r.latency for r in reqs)
        self.stats = {
            "requests": len(reqs), "wall_s": wall, "fovs_per_s": len(reqs) / wall if wall > 0 else None,
            "p50_latency_s": lat[len(lat) // 2] if lat else None,
            "p99_latency_s": lat[min(len(lat) - 1, int(np.ceil(0.99 * len(lat))) - 1)] if lat else None,
            "launches": launches, "tiles": tiles, "overflow_batches": overflows, "batch_tiles": self.B,
            "slots": len(self.slots),
            "hbm_budget_bytes": self.budget,
        }
        return reqs


def poisson_arrivals(n, rate, seed=0):
    """``n`` arrival times (s) of a Poisson process of ``rate`` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate > 0 else np.zeros(n)
