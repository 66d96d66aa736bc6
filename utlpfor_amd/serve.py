"""Serving loop over a compressed store resident in HBM (continuous batching).

The reference is a library of per-block calls; a server built on it answers
each request with a loop of UnpackUint32 / PackUint32 / GetUint32 calls
(unpack.go:5, pack.go:76, get.go:74). Here requests are batched instead:

* Store: many streams (uint32 and uint64, each a concatenation of blocks
  exactly as PackUint32(flag|Append) builds it) laid end to end in one HBM
  buffer, with one global index (block byte offsets and value offsets from
  the device scan of the encoders' per-block meta words).
* Requests: decompress a block range of a stream, compress a value array,
  or Get up to 1024 positions of a stream.
* Engine: each tick admits the requests that have arrived, in arrival
  order, while their device memory fits what the device has free: the
  tick's heap half, which grows (up to the device's free memory at that
  moment, less a reserve) when the admitted requests need more. Then it runs
  the whole batch with one launch per kind and type: a gathered index
  (utl_gather_index) plus one decode for all decompress requests, one
  batched Get, one batched encode. Requests that arrive meanwhile join the
  next tick. A request's latency is from its arrival to the end of the tick
  that answers it, including the copy of its result to the host.

Every result equals what the request gets when it runs alone (the kernels
treat blocks independently); tests check that bit for bit.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib, batch
from .device import Context, DeviceArray, PinnedBuffer, check, mem_info

DECOMPRESS, COMPRESS, GET = 0, 1, 2
KIND_NAMES = ("decompress", "compress", "get")
_ALIGN = 256


@dataclass
class StoreStream:
    first_block: int   # global block index of the stream's first block
    nblocks: int
    nvalues: int
    is64: bool
    source: int        # index of the source array the stream encodes
    flag: int


class Store:
    """Streams resident in HBM with one global index. Each source array is
    encoded once on the GPU; the store holds `replicas` copies of every
    encoded source (device-to-device copies), or as many as fill
    `target_bytes` of compressed data."""

    def __init__(self, ctx: Context, sources: list[np.ndarray], flags: list[int], replicas: int = 1,
                 target_bytes: int | None = None):
        lib = _lib.load()
        self.ctx = ctx
        self.sources = sources
        enc = [batch.encode(ctx, s, flag=f) for s, f in zip(sources, flags)]
        per_round = sum(e.nbytes for e in enc)
        if target_bytes is not None:
            replicas = max(1, int(target_bytes // max(per_round, 1)))
        self.replicas = replicas
        nbytes = per_round * replicas
        nblocks = sum(e.nblocks for e in enc) * replicas
        self.data = DeviceArray(nbytes + 64, np.uint8, ctx)
        self.meta = DeviceArray(max(nblocks, 1), np.uint32, ctx)
        self.streams: list[StoreStream] = []
        boff = blk = 0
        for _ in range(replicas):
            for i, e in enumerate(enc):
                check(lib.utl_dev_copy(ctx.handle, self.data.ptr + boff, e.data.ptr, e.nbytes), "store copy")
                check(lib.utl_dev_copy(ctx.handle, self.meta.ptr + 4 * blk, e.meta.ptr, 4 * e.nblocks), "meta copy")
                self.streams.append(StoreStream(blk, e.nblocks, e.nvalues, e.is64, i, flags[i]))
                boff += e.nbytes
                blk += e.nblocks
        self.nbytes, self.nblocks = nbytes, nblocks
        self.block_off, self.out_off = batch.build_index(ctx, self.meta, nblocks)
        ctx.sync()
        for e in enc:
            e.free()

    def expected(self, sid: int, v0: int, v1: int) -> np.ndarray:
        """Source values [v0, v1) of stream sid (for checks)."""
        return self.sources[self.streams[sid].source][v0:v1]

    def free(self) -> None:
        for a in (self.data, self.meta, self.block_off, self.out_off):
            a.free()


@dataclass
class Request:
    kind: int
    arrival: float = 0.0            # seconds after the start of serving
    stream: int = -1                # decompress / get
    first: int = 0                  # decompress: first block within the stream
    nblocks: int = 0                # decompress: blocks
    positions: np.ndarray | None = None   # get: value positions within the stream
    values: np.ndarray | None = None      # compress: uint32 or uint64 values
    flag: int = 0                   # compress
    # set by the engine
    done: float = math.nan
    tick: int = -1
    result: np.ndarray | None = None      # values, Get values, or packed bytes
    status: np.ndarray | None = None      # per-block (decompress) / per-position (get) status words
    block_len: np.ndarray | None = None   # compress: byte length of each packed block

    def latency(self) -> float:
        return self.done - self.arrival


class _Heap:
    """Bump allocator over one device allocation, reset every tick."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.mem = DeviceArray(nbytes, np.uint8, ctx)
        self.cap = nbytes
        self.top = 0

    def grow(self, nbytes: int) -> None:
        """Replaces the allocation by one of nbytes (its contents are not
        kept: only between ticks)."""
        if nbytes > self.cap:
            self.mem.free()
            self.mem = DeviceArray(nbytes, np.uint8, self.ctx)
            self.cap = nbytes

    def reset(self) -> None:
        self.top = 0

    def take(self, nbytes: int) -> int:
        off = (self.top + _ALIGN - 1) // _ALIGN * _ALIGN
        if off + nbytes > self.cap:
            raise MemoryError("serving heap exhausted")
        self.top = off + nbytes
        return self.mem.ptr + off


class _HostStage:
    """Copies between the device and request arrays in host memory through a
    pinned buffer (full link rate) in chunks, each copied out to (or in
    from) the arrays it holds by a pool of threads (numpy copies release the
    GIL; one host thread copies at about 14 GB/s). The buffer's two halves
    take chunks in turn, so the link's copy of one chunk overlaps the host
    copies of the next."""

    def __init__(self, ctx: Context, nbytes: int = 256 << 20, threads: int = 8):
        self.ctx, self.lib = ctx, _lib.load()
        self.pin = PinnedBuffer(nbytes)
        self.pool = ThreadPoolExecutor(max_workers=threads)
        self.mover = ThreadPoolExecutor(max_workers=1)  # the link's copies, one at a time
        self.piece = 4 << 20  # bytes one thread copies at a time

    def fetch(self, dev: int, nbytes: int, dsts: list[tuple[int, np.ndarray]]) -> None:
        """Copies device bytes [dev + off, dev + off + a.nbytes) into array a
        for every (off, a) in dsts (contiguous arrays, disjoint ranges inside
        [0, nbytes)), through the two halves of the pinned buffer in turn:
        the next chunk comes over into one half (on its own thread) while
        the host threads copy the current one out of the other."""
        dsts = sorted(((off, a.reshape(-1).view(np.uint8)) for off, a in dsts if a.nbytes), key=lambda t: t[0])
        if not dsts or nbytes <= 0:
            return
        cap = max(1, self.pin.nbytes // 2)
        end = min(nbytes, max(off + a.nbytes for off, a in dsts))  # nothing past the last array comes over
        chunks = [(c0, min(c0 + cap, end)) for c0 in range(0, end, cap)]
        k = 0
        nxt = None
        if len(chunks) == 1:  # one chunk: nothing to overlap, no thread hop
            self._download(self.pin.ptr, dev, chunks[0][1])
        else:
            nxt = self.mover.submit(self._download, self.pin.ptr, dev, chunks[0][1])
        try:
            for n, (c0, c1) in enumerate(chunks):
                if nxt is not None:
                    nxt.result()
                nxt = None
                base = (n & 1) * cap
                if n + 1 < len(chunks):  # the other half's copies out ended in the last round
                    b0, b1 = chunks[n + 1]
                    nxt = self.mover.submit(self._download, self.pin.ptr + ((n + 1) & 1) * cap, dev + b0, b1 - b0)
                jobs = []
                j = k
                while j < len(dsts) and dsts[j][0] < c1:
                    off, a = dsts[j]
                    lo, hi = max(off, c0), min(off + a.nbytes, c1)
                    for x in range(lo, hi, self.piece):
                        y = min(x + self.piece, hi)
                        src, dst = self.pin.bytes[base + x - c0:base + y - c0], a[x - off:y - off]
                        if hi - lo <= self.piece:
                            dst[:] = src
                        else:
                            jobs.append(self.pool.submit(np.copyto, dst, src))
                    if off + a.nbytes <= c1:
                        k = j + 1
                    j += 1
                for f in jobs:
                    f.result()
        finally:
            if nxt is not None:
                nxt.exception()  # the download ends before the buffer is reused

    def _download(self, host: int, dev: int, nbytes: int) -> None:
        check(self.lib.utl_memcpy_d2h(self.ctx.handle, host, dev, nbytes), "d2h")

    def put(self, dev: int, nbytes: int, srcs: list[tuple[int, np.ndarray]]) -> None:
        """The other direction: array a's bytes to device [dev + off, ...)
        for every (off, a) in srcs (contiguous arrays, disjoint ranges that
        cover [0, nbytes)), through the two halves of the pinned buffer in
        turn: the host threads fill one half while the other half's upload
        runs (on its own thread; the upload call waits for its copy)."""
        srcs = sorted(((off, a.reshape(-1).view(np.uint8)) for off, a in srcs if a.nbytes), key=lambda t: t[0])
        cap = max(1, self.pin.nbytes // 2)
        c0, k, n = 0, 0, 0
        uploads = [None, None]  # the upload under way from each half
        try:
            while c0 < nbytes:
                c1 = min(c0 + cap, nbytes)
                h = n & 1
                if uploads[h] is not None:
                    uploads[h].result()  # the half is free again
                    uploads[h] = None
                base = h * cap
                jobs = []
                j = k
                while j < len(srcs) and srcs[j][0] < c1:
                    off, a = srcs[j]
                    lo, hi = max(off, c0), min(off + a.nbytes, c1)
                    for x in range(lo, hi, self.piece):
                        y = min(x + self.piece, hi)
                        dst, src = self.pin.bytes[base + x - c0:base + y - c0], a[x - off:y - off]
                        if hi - lo <= self.piece:
                            dst[:] = src
                        else:
                            jobs.append(self.pool.submit(np.copyto, dst, src))
                    if off + a.nbytes <= c1:
                        k = j + 1
                    j += 1
                for f in jobs:
                    f.result()
                if c0 == 0 and c1 == nbytes:  # one chunk: nothing to overlap, no thread hop
                    self._upload(dev, self.pin.ptr + base, c1)
                else:
                    uploads[h] = self.mover.submit(self._upload, dev + c0, self.pin.ptr + base, c1 - c0)
                c0, n = c1, n + 1
        finally:
            for f in uploads:
                if f is not None:
                    f.exception()  # every upload ends before the buffer is reused or the caller goes on
        for f in uploads:
            if f is not None:
                f.result()

    def _upload(self, dev: int, host: int, nbytes: int) -> None:
        check(self.lib.utl_memcpy_h2d(self.ctx.handle, dev, host, nbytes), "h2d")

    def close(self) -> None:
        self.pool.shutdown()
        self.mover.shutdown()
        self.pin.free()


def _values_of(st: StoreStream, first: int, nblocks: int) -> int:
    return max(0, min(nblocks * 128, st.nvalues - first * 128))


class Engine:
    """Continuous-batching executor over a Store. Kernels run on ctx's
    stream; results go to the host through a second context (its own
    stream and pinned buffer), so that serve() delivers a tick's results
    while the next tick's kernels run. The heap is two halves that ticks
    use in turn: a half is reused only once its tick's results are
    delivered."""

    def __init__(self, ctx: Context, store: Store, heap_bytes: int | None = None, heap_fraction: float = 0.05,
                 reserve_fraction: float = 0.05, grow: bool = True):
        """heap_bytes (or heap_fraction of the free device memory) is the
        heap's starting size; with grow, a tick's half grows while the
        device keeps reserve_fraction of its memory free."""
        free, total = mem_info()
        if heap_bytes is None:
            heap_bytes = int(free * heap_fraction)
        self.ctx, self.store = ctx, store
        self.lib = _lib.load()
        half = heap_bytes // 2
        self.heaps = [_Heap(ctx, half), _Heap(ctx, half)]
        self.heap = self.heaps[0]
        self.grow = grow
        self.reserve = int(total * reserve_fraction)
        self.growths = 0
        self.stage = _HostStage(ctx)  # uploads, on ctx's stream
        self.dctx = Context(getattr(ctx, "device", 0))
        self.dstage = _HostStage(self.dctx)  # deliveries, on the second stream

    def close(self) -> None:
        """Frees the heaps and the host staging buffers."""
        self.stage.close()
        self.dstage.close()
        self.dctx.close()
        for h in self.heaps:
            h.mem.free()

    # ---- admission ----
    def admission_limit(self, half: int) -> int:
        """Device bytes a tick on heap half `half` may take: its size, or,
        when the heap may grow, its size plus the device's free memory now,
        less the reserve (another process or context may hold the rest)."""
        cap = self.heaps[half].cap
        if not self.grow:
            return cap
        free, _ = mem_info()
        return cap + max(0, free - self.reserve)

    def fit_half(self, half: int, need: int) -> None:
        """Grows heap half `half` to hold need bytes (between ticks)."""
        h = self.heaps[half]
        if need > h.cap:
            h.grow(max(need, min(2 * h.cap, self.admission_limit(half))))
            self.growths += 1

    def cost(self, r: Request) -> int:
        """Device bytes the request takes in a tick (with alignment slack)."""
        if r.kind == DECOMPRESS:
            st = self.store.streams[r.stream]
            n = r.nblocks
            return _values_of(st, r.first, n) * (8 if st.is64 else 4) + (n + 1) * 16 + (n + 1) * 4 + 32 + 4 * _ALIGN
        if r.kind == GET:
            return r.positions.size * (8 + 8 + 8 + 4) + 4 * _ALIGN
        v = r.values
        nb = (v.size + 127) // 128
        longest = batch.MAX_BLOCK64 if v.dtype == np.uint64 else batch.MAX_BLOCK32
        return v.nbytes + nb * (longest + 8 + 8 + 4 + 4 + 16) + 64 + 12 * _ALIGN

    def check_request(self, r: Request, limit: int | None = None) -> None:
        """Rejects a request outside its stream, or larger than the device
        memory the server may use (limit: admission_limit(0) by default)."""
        if r.kind == DECOMPRESS:
            st = self.store.streams[r.stream]
            if r.nblocks < 1 or r.first < 0 or r.first + r.nblocks > st.nblocks:
                raise ValueError("decompress range outside the stream")
        elif r.kind == GET:
            st = self.store.streams[r.stream]
            p = r.positions
            if p.size == 0 or p.min() < 0 or p.max() >= st.nvalues:
                raise ValueError("get position outside the stream")
        elif r.values.size == 0:
            raise ValueError("compress of no values")
        if self.cost(r) > (self.admission_limit(0) if limit is None else limit):
            raise ValueError("request larger than the device memory the server may use")

    # ---- one tick ----
    def run_batch(self, reqs: list[Request], deliver: bool = True) -> None:
        """Runs one batch to completion (results copied to the host when
        deliver is set)."""
        pending = self._launch(reqs, 0)
        if deliver:
            for fn in pending:
                fn()

    def _launch(self, reqs: list[Request], half: int) -> list:
        """Launches one batch on heap half `half` and waits for its kernels;
        returns the functions that deliver its results (they use only the
        delivery context)."""
        ctx, store = self.ctx, self.store
        self.heap = heap = self.heaps[half]
        heap.reset()
        pending = []  # (request, kind-specific handles) to deliver after the sync
        for is64 in (False, True):
            dec = [r for r in reqs if r.kind == DECOMPRESS and store.streams[r.stream].is64 == is64]
            if dec:
                pending += self._decompress(dec, is64)
            gets = [r for r in reqs if r.kind == GET and store.streams[r.stream].is64 == is64]
            if gets:
                pending += self._get(gets, is64)
            comp = [r for r in reqs if r.kind == COMPRESS and (r.values.dtype == np.uint64) == is64]
            if comp:
                pending += self._compress(comp, is64)
        ctx.sync()
        return pending

    def _h2d(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = self.heap.take(max(a.nbytes, 1))
        if a.nbytes:
            check(self.lib.utl_memcpy_h2d(self.ctx.handle, p, a.ctypes.data, a.nbytes), "h2d")
        return p

    def _d2h(self, p: int, n: int, dtype) -> np.ndarray:
        """Device to host through the delivery context."""
        out = np.empty(n, dtype=dtype)
        if out.nbytes:
            check(self.lib.utl_memcpy_d2h(self.dctx.handle, out.ctypes.data, p, out.nbytes), "d2h")
        return out

    def _decompress(self, reqs: list[Request], is64: bool):
        lib, ctx, store = self.lib, self.ctx, self.store
        rows = np.empty((len(reqs), 4), dtype=np.uint64)
        u = e = 0
        spans = []
        for i, r in enumerate(reqs):
            st = store.streams[r.stream]
            nv = _values_of(st, r.first, r.nblocks)
            rows[i] = (st.first_block + r.first, r.nblocks, u, e)
            spans.append((u, nv, e, r.nblocks))
            u += nv
            e += r.nblocks + 1
        item = 8 if is64 else 4
        d_req = self._h2d(rows.reshape(-1))
        gboff = self.heap.take(8 * e)
        gooff = self.heap.take(8 * e)
        out = self.heap.take(max(u * item, 1))
        status = self.heap.take(4 * max(e - 1, 1))
        check(lib.utl_gather_index(ctx.handle, store.block_off.ptr, store.out_off.ptr, store.nblocks, d_req,
                                   len(reqs), e, gboff, gooff), "utl_gather_index")
        f = lib.utl_decode_u64 if is64 else lib.utl_decode_u32
        check(f(ctx.handle, store.data.ptr, gboff, gooff, e - 1, out, status, None), "gathered decode")

        def deliver():
            dt = np.uint64 if is64 else np.uint32
            for r, (u0, nv, e0, nb) in zip(reqs, spans):
                r.result = np.empty(nv, dt)
                r.status = np.empty(nb, np.int32)
            self.dstage.fetch(out, u * item, [(u0 * item, r.result) for r, (u0, nv, e0, nb) in zip(reqs, spans)])
            self.dstage.fetch(status, 4 * max(e - 1, 0), [(4 * e0, r.status) for r, (u0, nv, e0, nb) in zip(reqs, spans)])
        return [deliver]

    def _get(self, reqs: list[Request], is64: bool):
        lib, ctx, store = self.lib, self.ctx, self.store
        blocks, pos, spans = [], [], []
        n = 0
        for r in reqs:
            st = store.streams[r.stream]
            p = r.positions.astype(np.int64)
            blocks.append((st.first_block + p // 128).astype(np.uint64))
            pos.append(p % 128)
            spans.append((n, p.size))
            n += p.size
        d_blk = self._h2d(np.concatenate(blocks))
        d_pos = self._h2d(np.concatenate(pos).astype(np.int64))
        out = self.heap.take(8 * n)
        status = self.heap.take(4 * n)
        check(lib.utl_get_batch(ctx.handle, int(is64), store.data.ptr, store.block_off.ptr, store.out_off.ptr,
                                store.nblocks, d_blk, d_pos, n, out, status), "utl_get_batch")

        def deliver():
            dt, item = (np.uint64, 8) if is64 else (np.uint32, 4)
            for r, (a, k) in zip(reqs, spans):
                r.result = np.empty(k, dt)
                r.status = np.empty(k, np.int32)
            self.dstage.fetch(out, item * n, [(item * a, r.result) for r, (a, k) in zip(reqs, spans)])
            self.dstage.fetch(status, 4 * n, [(4 * a, r.status) for r, (a, k) in zip(reqs, spans)])
        return [deliver]

    def _compress(self, reqs: list[Request], is64: bool):
        lib, ctx = self.lib, self.ctx
        dtype = np.uint64 if is64 else np.uint32
        lens = np.array([r.values.size for r in reqs], dtype=np.uint64)
        blk_off = batch.plan_blocks(lens)
        nb = int(blk_off[-1])
        val_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        item = np.dtype(dtype).itemsize
        d_vals = self.heap.take(max(int(val_off[-1]) * item, 1))
        self.stage.put(d_vals, int(val_off[-1]) * item,
                       [(int(val_off[i]) * item, np.ascontiguousarray(r.values, dtype=dtype)) for i, r in enumerate(reqs)])
        d_val_off = self._h2d(val_off)
        d_blk_off = self._h2d(blk_off)
        d_flags = self._h2d(np.array([r.flag for r in reqs], dtype=np.uint8))
        in_off = self.heap.take(8 * nb)
        in_meta = self.heap.take(4 * nb)
        check(lib.utl_make_jobs(ctx.handle, d_val_off, d_blk_off, d_flags, len(reqs), in_off, in_meta),
              "utl_make_jobs")
        # the sized form: lengths (and for uint32 each block's plan), the
        # index scan, then every block written at its offset
        meta = self.heap.take(4 * nb)
        boff = self.heap.take(8 * (nb + 1))
        ooff = self.heap.take(8 * (nb + 1))
        cap = nb * (batch.MAX_BLOCK64 if is64 else batch.MAX_BLOCK32) + 64
        data = self.heap.take(cap)
        if is64:
            check(lib.utl_encode_u64_size(ctx.handle, d_vals, in_off, in_meta, nb, meta), "utl_encode_u64_size")
            check(lib.utl_index_build(ctx.handle, meta, nb, boff, ooff), "utl_index_build")
            check(lib.utl_encode_u64_write(ctx.handle, d_vals, in_off, in_meta, nb, boff, data, cap),
                  "utl_encode_u64_write")
        else:
            plan = self.heap.take(8 * nb)
            check(lib.utl_encode_u32_size(ctx.handle, d_vals, in_off, in_meta, nb, meta, plan), "utl_encode_u32_size")
            check(lib.utl_index_build(ctx.handle, meta, nb, boff, ooff), "utl_index_build")
            check(lib.utl_encode_u32_write(ctx.handle, d_vals, in_off, in_meta, nb, plan, boff, data, cap),
                  "utl_encode_u32_write")

        def deliver():
            bo = self._d2h(boff, nb + 1, np.uint64)
            metas = self._d2h(meta, nb, np.uint32)
            dsts = []
            for i, r in enumerate(reqs):
                b0, b1 = int(blk_off[i]), int(blk_off[i + 1])
                r.result = np.empty(int(bo[b1]) - int(bo[b0]), np.uint8)
                r.block_len = (metas[b0:b1] & 0xFFFFFF).astype(np.int64)
                dsts.append((int(bo[b0]), r.result))
            self.dstage.fetch(data, int(bo[-1]), dsts)
        return [deliver]

    # ---- the loop ----
    def serve(self, requests: list[Request], deliver: bool = True, clock=time.perf_counter) -> dict:
        """Serves requests by arrival time (seconds from the call). Returns
        latency and throughput statistics."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        limit0 = self.admission_limit(0)
        for r in reqs:
            self.check_request(r, limit0)
        t0 = clock()
        i = tick = 0
        sizes = []
        worker = ThreadPoolExecutor(max_workers=1) if deliver else None
        inflight = [None, None]  # delivery of the last tick on each heap half

        def finish(pending, b, tick):
            for fn in pending:
                fn()
            t = clock() - t0
            for r in b:
                r.done, r.tick = t, tick

        try:
            while i < len(reqs):
                now = clock() - t0
                if reqs[i].arrival > now:
                    time.sleep(min(reqs[i].arrival - now, 0.001))
                    continue
                half = tick & 1
                if inflight[half] is not None:
                    inflight[half].result()  # the half's results are on the host: reuse it
                    inflight[half] = None
                limit = self.admission_limit(half)
                used, b = 0, []
                while i < len(reqs) and reqs[i].arrival <= now:
                    c = self.cost(reqs[i])
                    if b and used + c > limit:
                        break
                    b.append(reqs[i])
                    used += c
                    i += 1
                self.fit_half(half, used)
                pending = self._launch(b, half)
                if worker is not None:
                    inflight[half] = worker.submit(finish, pending, b, tick)
                else:
                    finish([], b, tick)
                tick += 1
                sizes.append(len(b))
        finally:
            # every delivery under way completes (or fails) before serve
            # returns or raises, so no later tick can reuse its heap half
            for f in inflight:
                if f is not None:
                    f.exception()
            if worker is not None:
                worker.shutdown(wait=True)
        for f in inflight:
            if f is not None:
                f.result()  # a delivery's own error
        wall = clock() - t0
        return summarize(reqs, wall, self.store, ticks=tick, mean_batch=float(np.mean(sizes)) if sizes else 0.0,
                         max_batch=int(max(sizes)) if sizes else 0, heap_growths=self.growths)


def request_bytes(r: Request, store: Store) -> int:
    """Bytes a request delivers (decoded values, Get values) or consumes
    (values to compress)."""
    if r.kind == DECOMPRESS:
        st = store.streams[r.stream]
        return _values_of(st, r.first, r.nblocks) * (8 if st.is64 else 4)
    if r.kind == GET:
        return r.positions.size * (8 if store.streams[r.stream].is64 else 4)
    return r.values.nbytes


def summarize(reqs: list[Request], wall: float, store: Store, **extra) -> dict:
    lat = np.array([r.latency() for r in reqs]) * 1e3
    out = {"requests": len(reqs), "wall_s": round(wall, 4),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "GBps": round(sum(request_bytes(r, store) for r in reqs) / wall / 1e9, 3)}
    for k, name in enumerate(KIND_NAMES):
        lk = np.array([r.latency() for r in reqs if r.kind == k]) * 1e3
        if lk.size:
            out[name] = {"n": int(lk.size), "p50_ms": round(float(np.percentile(lk, 50)), 3),
                         "p99_ms": round(float(np.percentile(lk, 99)), 3)}
    out.update(extra)
    return out


def make_requests(store: Store, pool: list[tuple[np.ndarray, int]], n: int, rate: float, seed: int = 5,
                  min_values: int = 128, max_values: int = 4 << 20, mix=(0.6, 0.25, 0.15),
                  max_get: int = 1024) -> list[Request]:
    """Poisson arrivals at `rate` per second; kinds by `mix` (decompress,
    compress, get); sizes log-uniform in [min_values, max_values] values.
    Decompress reads a random block range of a random stream; Get reads
    1..max_get random positions; compress takes a random slice of a random
    (values, flag) pool entry."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    kinds = rng.choice(3, n, p=np.asarray(mix) / np.sum(mix))
    sizes = np.floor(np.exp(rng.uniform(np.log(min_values), np.log(max_values + 1), n))).astype(np.int64)
    reqs = []
    for t, k, sz in zip(arrivals, kinds, sizes):
        if k == DECOMPRESS:
            sid = int(rng.integers(len(store.streams)))
            st = store.streams[sid]
            nb = int(min(max(1, (sz + 127) // 128), st.nblocks))
            first = int(rng.integers(0, st.nblocks - nb + 1))
            reqs.append(Request(DECOMPRESS, float(t), stream=sid, first=first, nblocks=nb))
        elif k == GET:
            sid = int(rng.integers(len(store.streams)))
            st = store.streams[sid]
            npos = int(rng.integers(1, max_get + 1))
            reqs.append(Request(GET, float(t), stream=sid, positions=rng.integers(0, st.nvalues, npos)))
        else:
            vals, flag = pool[int(rng.integers(len(pool)))]
            m = int(min(sz, vals.size))
            a = int(rng.integers(0, vals.size - m + 1))
            reqs.append(Request(COMPRESS, float(t), values=vals[a:a + m], flag=flag))
    return reqs
