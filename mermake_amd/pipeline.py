"""The hyb and DAPI callers of the hot path (reference ``mermake/__main__.py:213-269``).

``HybPipeline`` replaces the per-channel hyb loop (``__main__.py:242-269``):

    for x, y, tile, raw in deconvolver.hybs.tile_wise(chan, flat, **hybs):
        Xh = find_local_maxima(tile, raw=raw, **hybs)
        keep = all(ov <= Xh[:, 1:3] < tile.shape[1:3] - ov); Xh = Xh[keep]
        Xh[:, 1] += x - ov; Xh[:, 2] += y - ov
    Xhf = vstack(...)

with one batch of tiles per launch: deconvolution, blur subtraction, maxima,
fit and keep-filter of every tile run as a handful of kernels, with a single
host synchronisation per channel stack (to read the number of spots).  The
``raw`` values the fit needs are evaluated from the stack and flat at the
sampled voxels instead of materialising the padded tile.

``plain=True`` drives the same kernels the way the reference drives its own:
one tile at a time, host syncs after the maxima and after the fit.

``DapiPipeline`` is ``__main__.py:217-225``: ``apply`` with the DAPI settings,
RMS normalisation, maxima of +buffer and of -buffer (both signs in one pass
over the volume; the negation is folded into the kernel).
"""
import numpy as np
import torch

from . import _lib as L
from ._dev import capture_graph
from .deconvolver import Deconvolver, prepare_flat, prepare_stack, tile_table
from .maxima import find_local_maxima_pm


class HybPipeline:
    def __init__(self, deconvolver, blur_radius=30, threshold=3600, delta=1, delta_fit=3, sigmaZ=1, sigmaXY=1.5,
                 cap=16384, plain=False, window=True, graph=False, spectral=True, **kwargs):
        if delta > 5:
            raise TypeError("Delta must be an less than or equal to 5 due to MAX_KERNEL_POINTS")
        self.d = deconvolver
        self.blur_radius = blur_radius
        self.threshold = float(torch.tensor(float(threshold), dtype=torch.float32))
        self.delta = int(delta)
        self.delta_fit = int(delta_fit)
        self.sigmaZ = float(sigmaZ)
        self.sigmaXY = float(sigmaXY)
        self.cap = int(cap)
        self.plain = plain
        self.window = window
        self.graph = bool(graph) and not plain
        # the blur subtraction rides on the Wiener operator when every box the windows read stays
        # inside the tile (Deconvolver.spectral_blur); otherwise the blur pass runs
        self.spectral = bool(spectral)
        self.prepare()
        self._tables = {}
        self._ws = {}
        self._graphs = {}
        self._cap_stream = None

    @classmethod
    def from_settings(cls, psfs, channel_shape, tile_size=500, overlap=89, beta=1e-4, zpad=None, psf_mode="zip",
                      autotune=False, **hybs):
        """Deconvolver with ``zpad = sz - 1`` as in ``__main__.py:184,195``; ``psf_mode`` pairs
        several PSFs with the tiles (``Deconvolver``).  ``autotune`` (True or the keyword arguments of
        ``autotune.tune``) times the deconvolution schedules on this channel shape, with the
        deconvolution window of the pipeline, and keeps the fastest."""
        if zpad is None:
            zpad = channel_shape[0] - 1
        d = Deconvolver(psfs, channel_shape, tile_size=tile_size, zpad=zpad, overlap=overlap, beta=beta,
                        psf_mode=psf_mode)
        pipe = cls(d, **hybs)
        if autotune:
            from .autotune import tune
            opts = dict(autotune) if isinstance(autotune, dict) else {}
            opts.setdefault("window", pipe.deconv_window())
            tune(d, channel_shape, **opts)
            pipe.prepare()
        return pipe

    def prepare(self):
        """Build the operators this pipeline reads (the box-folded ones of a spectral blur) now
        rather than inside the first batch."""
        if self.spectral_blur():
            self.d.box_spectra(self.blur_radius)

    def deconv_window(self):
        """Rows and columns [lo, hi) of every tile that the deconvolution computes for this pipeline
        (the blur window widened by the box half-width), or None for whole tiles."""
        win = self.windows()
        if win is None:
            return None
        if self.spectral_blur():
            return win[0]
        T = self.d._plan.T
        delta = 0 if self.blur_radius is None else int(self.blur_radius) // 2
        return max(0, win[0][0] - delta), min(T, win[0][1] + delta)

    def spectral_blur(self):
        """Whether the blur subtraction is folded into the Wiener operator for this pipeline."""
        win = self.windows()
        return self.spectral and win is not None and self.d.spectral_blur(self.blur_radius, win[0])

    def _buf(self, name, shape, dtype):
        t = self._ws.get(name)
        n = 1
        for s in shape:
            n *= s
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(n, dtype=dtype, device=self.d.device)
            self._ws[name] = t
        return t[:n].view(shape)

    def _table(self, img, flat):
        key = (img.data_ptr(), tuple(img.shape), tuple(img.stride()), img.dtype,
               None if flat is None else flat.data_ptr())
        t = self._tables.get(key)
        if t is None:
            geom = self.d.geometry(img.shape)
            t = (geom, tile_table(geom, img, flat))
            if len(self._tables) > 8:
                self._tables.clear()
            self._tables[key] = t
        return t

    def windows(self):
        """(blur window, candidate window) as (lo, hi) pairs over x and y, or None for whole tiles.

        The keep filter drops every spot whose fitted x or y leaves [ov, extent - ov), and the fit
        moves a candidate by at most delta_fit (the centroid is a weighted mean of sphere offsets
        with weights >= 0), so candidates outside [ov - delta_fit - 1, T - ov + delta_fit + 1) are
        never kept.  Their neighbour tests and fits read the blurred tile up to max(delta,
        delta_fit) further out, which is where the blur has to be computed."""
        if not self.window:
            return None
        d = self.d
        T, ov = d._plan.T, d.overlap
        c0, c1 = ov - self.delta_fit - 1, T - ov + self.delta_fit + 1
        r = max(self.delta, self.delta_fit)
        b0, b1 = c0 - r, c1 + r
        if b0 <= 0 or b1 >= T or c0 >= c1:
            return None
        return (b0, b1), (c0, c1)

    def _spots_batch(self, table, ntiles, single, cap):
        """Enqueue deconv -> blur -> maxima -> fit -> keep for ``ntiles`` descriptors."""
        d = self.d
        T, sz = d._plan.T, d.tile_height
        st = L.stream_ptr(d.device)
        win = self.windows()
        ws_bytes = L.lib().mm_maxima_ws_bytes(sz, T, T, ntiles)
        ws = self._buf("mws", (ws_bytes,), torch.uint8)
        coords = self._buf("coords", (ntiles, cap, 3), torch.int32)
        counts = self._buf("counts", (ntiles,), torch.int32)
        fit = self._buf("fit", (ntiles, cap, 8), torch.float32)
        # the blur subtraction rides on the operator, or else the threshold test of the maxima rides
        # on the blur pass (one entry) where the blur width allows
        spectral = self.spectral_blur()
        fused = win is not None and (spectral or d.fused_blur_maxima(self.blur_radius))
        mx = (self.threshold, self.delta, win[1], coords, cap, counts, ws, ws_bytes) if fused else None
        res = self._buf("res", (ntiles, sz, T, T), torch.float32)
        res = d.run_tiles(table, ntiles, ntiles if single else 1, self.blur_radius, res_out=res,
                          window=None if win is None else win[0], maxima=mx, spectral=spectral)
        if win is None:
            L.call("mm_tiles_maxima", d._plan.h, L.ptr(table), ntiles, L.ptr(res), self.threshold, self.delta,
                   L.ptr(coords), cap, L.ptr(counts), L.ptr(ws), ws_bytes, st)
        elif not fused:
            L.call("mm_tiles_maxima_window", d._plan.h, L.ptr(table), ntiles, L.ptr(res), self.threshold,
                   self.delta, win[1][0], win[1][1], L.ptr(coords), cap, L.ptr(counts), L.ptr(ws), ws_bytes, st)
        L.call("mm_tiles_fit", d._plan.h, L.ptr(table), ntiles, L.ptr(res), L.ptr(coords), L.ptr(counts), cap,
               self.delta_fit, self.sigmaZ, self.sigmaXY, L.ptr(fit), st)
        return counts, fit

    def tile_volumes(self, stack, flat_field=None, fill=-1e30):
        """Host copies ``(x, y, volume, raw)`` of what the maxima and the fit of each tile read, over the
        tile's view ``(sz, xend, yend)`` (``deconvolver.py:170``): the deconvolved tile minus its blur as
        this pipeline computes it (through the operator or by the blur pass), with ``fill`` outside the
        blur window, which nothing reads, and the raw view.  The oracle's maxima, fit and keep filter on
        these volumes give this pipeline's kept rows; tests use it for that."""
        d = self.d
        d._check(stack)
        img = prepare_stack(stack)
        flat = prepare_flat(flat_field, img)
        geom, table = self._table(img, flat)
        T, sz = d._plan.T, d.tile_height
        n = len(geom)
        single = d.psf_fft.shape[0] == 1
        win = self.windows()
        res = self._buf("res", (n, sz, T, T), torch.float32)
        vol = d.run_tiles(table, n, n if single else 1, self.blur_radius, res_out=res,
                          window=None if win is None else win[0], spectral=self.spectral_blur())
        vol = vol.cpu().numpy()
        raw = torch.empty((n, sz, T, T), dtype=torch.float32, device=d.device)
        L.call("mm_deconv_tile_raw", d._plan.h, L.ptr(table), n, L.ptr(raw), L.stream_ptr(d.device))
        raw = raw.cpu().numpy()
        out = []
        for i, t in enumerate(geom):
            v = vol[i, :, :t["xe"], :t["ye"]].copy()
            if win is not None:
                lo, hi = win[0]
                m = np.full(v.shape[1:], True)
                m[lo:hi, lo:hi] = False
                v[:, m] = fill
            out.append((t["x"], t["y"], v, raw[i, :, :t["xe"], :t["ye"]].copy()))
        return out

    def _keep_launch(self, table, n, counts, fit, cap, xh, stats):
        d = self.d
        L.call("mm_tiles_keep", L.ptr(table), n, d.overlap, L.ptr(fit), L.ptr(counts), cap, L.ptr(stats[2:]),
               L.ptr(xh), n * cap, L.ptr(stats[1:2]), L.ptr(stats[0:1]), L.stream_ptr(d.device))

    def _keep(self, table, n, counts, fit, cap):
        """Keep-filter of a batch: the kept rows (device) and, on the host, ``[largest candidate
        count, rows kept, kept per tile...]`` (one synchronisation)."""
        stats = self._buf("stats", (n + 2,), torch.int32)
        xh = torch.empty((n * cap, 8), dtype=torch.float32, device=self.d.device)
        self._keep_launch(table, n, counts, fit, cap, xh, stats)
        return xh, stats.cpu()

    def _graph_batch(self, table, n, single, cap):
        """``_spots_batch`` + ``_keep`` as the replay of a hipGraph of the batch's kernel chain, one graph
        per (tiles, cap).  A graph owns its buffers and a table of ``n`` descriptors that each batch is
        copied into; the first batch of a shape runs eagerly (allocating those buffers and the library's
        lazy tables), and the graph is captured after it.  Returns the kept rows (a copy, since the
        graph's buffers are reused) and the host statistics of ``_keep``."""
        d = self.d
        key = (n, cap, single, d.generation)
        ent = self._graphs.get(key)
        if ent is not None:
            g, tbuf, xh, stats, _ = ent
            tbuf.copy_(table)
            g.replay()
            st = stats.cpu()
            return xh[:int(st[1])].clone(), st
        saved = (self._ws, d._ws)
        bufs = ({}, {})
        self._ws, d._ws = bufs
        try:
            tbuf = torch.empty_like(table)
            tbuf.copy_(table)
            stats = self._buf("stats", (n + 2,), torch.int32)
            xh = self._buf("xh", (n * cap, 8), torch.float32)
            counts, fit = self._spots_batch(tbuf, n, single, cap)
            self._keep_launch(tbuf, n, counts, fit, cap, xh, stats)
            st = stats.cpu()
            rows = xh[:int(st[1])].clone()
            if self._cap_stream is None:
                self._cap_stream = torch.cuda.Stream(d.device)

            def chain():
                counts, fit = self._spots_batch(tbuf, n, single, cap)
                self._keep_launch(tbuf, n, counts, fit, cap, xh, stats)
            g = capture_graph(chain, self._cap_stream)
        finally:
            self._ws, d._ws = saved
        for k in [k for k in self._graphs if k[3] != d.generation]:
            del self._graphs[k]     # chains of an older schedule hold the address of its operators
        while len(self._graphs) >= 4:
            self._graphs.pop(next(iter(self._graphs)))
        self._graphs[key] = (g, tbuf, xh, stats, bufs)
        return rows, st

    def tile_rows(self, table, n):
        """Kept spot rows of each of ``n`` tile descriptors (a device table), as host arrays;
        the candidate cap grows until no tile overflows it."""
        d = self.d
        single = d.psf_fft.shape[0] == 1
        cap = self.cap
        while True:
            counts, fit = self._spots_batch(table, n, single, cap)
            xh, stats = self._keep(table, n, counts, fit, cap)
            if int(stats[0]) > cap:
                cap = int(stats[0])
                continue
            k = stats[2:].numpy()
            rows = xh[:int(k.sum())].cpu().numpy()
            offs = np.concatenate([[0], np.cumsum(k)])
            return [rows[offs[i]:offs[i + 1]].copy() for i in range(n)]

    @staticmethod
    def _split(parts, owner, b0, nb, kept_h, xh):
        """Rows of a batch to their stacks: tiles of one stack are contiguous."""
        off = 0
        for i in range(nb):
            si = owner[b0 + i]
            if kept_h[i]:
                if parts[si] and parts[si][-1][2] is xh and parts[si][-1][1] == off:
                    parts[si][-1][1] = off + kept_h[i]
                else:
                    parts[si].append([off, off + kept_h[i], xh])
            off += kept_h[i]

    def __call__(self, stack, flat_field=None):
        """Spots of one channel stack: ``(N, 8)`` float32 rows in tile order, then raster order."""
        return self.run_many([stack], [flat_field])[0]

    def submit(self, stacks, flat_fields=None):
        """``run_many`` without waiting: the batch's kernel chain is enqueued on one of two side streams, each
        with its own workspaces, and a ``PendingSpots`` is returned; its ``result()`` waits and gives the
        ``run_many`` output.  A caller keeps up to two batches in flight, so the host work between batches
        and the last small kernels of one batch overlap the next batch's passes.  A third ``submit`` waits for
        the oldest batch first.  Stacks whose tiles do not fit one batch, graph pipelines and plain pipelines
        run ``run_many`` at once (the returned result is then already complete)."""
        d = self.d
        if flat_fields is None:
            flat_fields = [None] * len(stacks)
        if self.plain or self.graph:
            return PendingSpots(done=self.run_many(stacks, flat_fields))
        tables, owner, alive = [], [], []
        for si, (stack, flat_field) in enumerate(zip(stacks, flat_fields)):
            d._check(stack)
            img = prepare_stack(stack)
            flat = prepare_flat(flat_field, img)
            alive.append((img, flat))
            geom, tab = self._table(img, flat)
            tables.append(tab)
            owner += [si] * len(geom)
        n = len(owner)
        if d.batch_size(n, self.cap * 48 + d.tile_height * d._plan.T ** 2 // 8) < n:
            return PendingSpots(done=self.run_many(stacks, flat_fields))
        if not hasattr(self, "_slots"):
            self._slots = [{"ws": {}, "dws": {}, "stream": torch.cuda.Stream(d.device), "pending": None}
                           for _ in range(2)]
            self._next = 0
        slot = self._slots[self._next]
        self._next ^= 1
        if slot["pending"] is not None:
            slot["pending"].result()      # its workspaces are about to be reused
        single = d.psf_fft.shape[0] == 1
        cap = self.cap
        cur = torch.cuda.current_stream(d.device)
        st = slot["stream"]
        st.wait_stream(cur)               # the inputs were made on the caller's stream
        saved = (self._ws, d._ws)
        self._ws, d._ws = slot["ws"], slot["dws"]
        try:
            with torch.cuda.stream(st):
                table = tables[0] if len(tables) == 1 else torch.cat(tables)
                counts, fit = self._spots_batch(table, n, single, cap)
                stats = self._buf("stats", (n + 2,), torch.int32)
                xh = torch.empty((n * cap, 8), dtype=torch.float32, device=d.device)
                self._keep_launch(table, n, counts, fit, cap, xh, stats)
                stats_h = slot.get("stats_h")
                if stats_h is None or stats_h.numel() < n + 2:
                    stats_h = slot["stats_h"] = torch.empty(n + 2, dtype=torch.int32, pin_memory=True)
                stats_h[:n + 2].copy_(stats, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(st)
        finally:
            self._ws, d._ws = saved
        p = PendingSpots(pipe=self, stacks=stacks, flats=flat_fields, alive=(alive, table), owner=owner, n=n,
                         cap=cap, xh=xh, stats_h=stats_h, event=ev, stream=st)
        slot["pending"] = p
        return p

    def run_many(self, stacks, flat_fields=None):
        """Spots of several channel stacks batched together: tiles of different stacks (of any x, y size,
        the same depth) share every launch.  One ``(N, 8)`` tensor per stack, equal to ``self(stack, flat)``."""
        d = self.d
        if flat_fields is None:
            flat_fields = [None] * len(stacks)
        tables, owner = [], []
        # the prepared stacks and flats (device copies of host inputs, dtype casts) stay referenced
        # until every batch that reads them through the tile tables has finished
        alive = []
        for si, (stack, flat_field) in enumerate(zip(stacks, flat_fields)):
            d._check(stack)
            img = prepare_stack(stack)
            flat = prepare_flat(flat_field, img)
            alive.append((img, flat))
            geom, tab = self._table(img, flat)
            tables.append(tab)
            owner += [si] * len(geom)
        n = len(owner)
        table_all = tables[0] if len(tables) == 1 else torch.cat(tables)
        single = d.psf_fft.shape[0] == 1
        parts = [[] for _ in stacks]
        B = 1 if self.plain else d.batch_size(n, self.cap * 48 + d.tile_height * d._plan.T ** 2 // 8)
        rec = L.TILE_DTYPE.itemsize
        for b0 in range(0, n, B):
            nb = min(B, n - b0)
            table = table_all[b0 * rec:(b0 + nb) * rec]
            cap = self.cap
            while True:
                if self.graph:
                    xh, stats = self._graph_batch(table, nb, single, cap)
                    mx = int(stats[0])
                    if mx > cap:
                        cap = mx
                        continue
                    kept_h = stats[2:].tolist()
                    self._split(parts, owner, b0, nb, kept_h, xh)
                    break
                counts, fit = self._spots_batch(table, nb, single, cap)
                if self.plain:
                    mx = int(counts.cpu().max())        # sync after the maxima, as maxima.py:46-50
                else:
                    mx = None
                if mx is not None and mx > cap:
                    cap = mx
                    continue
                xh, stats = self._keep(table, nb, counts, fit, cap)   # one sync per batch
                mx = int(stats[0])
                if mx > cap:
                    cap = mx
                    continue
                kept_h = stats[2:].tolist()
                self._split(parts, owner, b0, nb, kept_h, xh)
                break
        del alive   # every batch has synchronised on its keep-filter statistics
        out = []
        for si in range(len(stacks)):
            views = [x[a:b] for a, b, x in parts[si]]
            if not views:
                out.append(torch.zeros((0, 8), dtype=torch.float32, device=d.device))
            else:
                out.append(views[0] if len(views) == 1 else torch.cat(views))
        return out


class PendingSpots:
    """Spots of a ``HybPipeline.submit`` batch; ``result()`` waits for them (once) and returns one ``(N, 8)``
    tensor per stack, equal to ``run_many`` of the same stacks."""

    def __init__(self, done=None, pipe=None, stacks=None, flats=None, alive=None, owner=None, n=0, cap=0, xh=None,
                 stats_h=None, event=None, stream=None):
        self._out = done
        self._pipe, self._stacks, self._flats, self._alive = pipe, stacks, flats, alive
        self._owner, self._n, self._cap, self._xh = owner, n, cap, xh
        self._stats_h, self._event, self._stream = stats_h, event, stream

    def result(self):
        if self._out is not None:
            return self._out
        self._event.synchronize()
        st = self._stats_h[:self._n + 2].clone()
        pipe = self._pipe
        if int(st[0]) > self._cap:        # a tile overflowed the candidate cap: run the batch again, larger
            out = pipe.run_many(self._stacks, self._flats)
        else:
            parts = [[] for _ in self._stacks]
            HybPipeline._split(parts, self._owner, 0, self._n, st[2:].tolist(), self._xh)
            out = []
            dev = pipe.d.device
            for si in range(len(self._stacks)):
                views = [x[a:b] for a, b, x in parts[si]]
                if not views:
                    out.append(torch.zeros((0, 8), dtype=torch.float32, device=dev))
                else:
                    out.append(views[0] if len(views) == 1 else torch.cat(views))
        # the rows were written on the side stream: later work on the caller's stream may use them, and their
        # memory goes back to the side stream's pool only after that work
        cur = torch.cuda.current_stream(pipe.d.device)
        cur.wait_stream(self._stream)
        if self._xh is not None:
            self._xh.record_stream(cur)
        self._out = out
        self._alive = self._stacks = self._flats = self._xh = None
        return out


class DapiPipeline:
    def __init__(self, deconvolver, blur_radius=50, threshold=3.0, delta=5, delta_fit=5, sigmaZ=1, sigmaXY=1.5,
                 spectral=True, **kwargs):
        if delta > 5:
            raise TypeError("Delta must be an less than or equal to 5 due to MAX_KERNEL_POINTS")
        self.d = deconvolver
        self.blur_radius = blur_radius
        self.kw = dict(threshold=float(torch.tensor(float(threshold), dtype=torch.float32)), delta=int(delta),
                       delta_fit=int(delta_fit), sigmaZ=float(sigmaZ), sigmaXY=float(sigmaXY))
        self._buffer = None
        # the blur subtraction rides on the Wiener operator where every box of the stitched window
        # stays inside the tile (Deconvolver.spectral_blur)
        self.spectral = bool(spectral)
        self.prepare()

    def _spectral(self):
        T, ov = self.d._plan.T, self.d.overlap
        return self.spectral and self.d.spectral_blur(self.blur_radius, (ov, T - ov))

    def prepare(self):
        """Build the box-folded operators of a spectral blur now rather than inside the first stack."""
        if self._spectral():
            self.d.box_spectra(self.blur_radius)

    def __call__(self, chan, flat_field=None):
        """``(Xh_plus, Xh_minus)`` of one DAPI stack."""
        d = self.d
        img = prepare_stack(chan)
        if self._buffer is None or tuple(self._buffer.shape) != tuple(img.shape):
            self._buffer = torch.empty(tuple(img.shape), dtype=torch.float32, device=d.device)
        buf = d.apply(img, flat_field=flat_field, blur_radius=self.blur_radius, output=self._buffer,
                      spectral=self.spectral)
        ws = torch.empty(L.lib().mm_rms_ws_bytes(), dtype=torch.uint8, device=d.device)
        rms = torch.empty(1, dtype=torch.float32, device=d.device)
        L.call("mm_rms_normalize", L.ptr(buf), buf.numel(), L.ptr(ws), L.ptr(rms), L.stream_ptr(d.device))
        # the maxima of +buffer and -buffer (__main__.py:224-225) from one pass over the volume
        return find_local_maxima_pm(buf, img, **self.kw)
