"""Serving loop on the GPU: every request's result in a mixed batch equals its
result alone, bit for bit, and both equal the oracle (decode values and
per-block statuses, Get values and statuses, packed bytes)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from utlpfor_amd.device import Context
    return Context(0)


def _store(ctx):
    from utlpfor_amd import serve, workloads
    sources = [workloads.plain_1m(20000, seed=11), workloads.delta_postings(30000, seed=12),
               workloads.timestamps64(10000, seed=13), workloads.uniform64(5000, seed=14)]
    return serve.Store(ctx, sources, [0, 1, 0, 0], replicas=2)


def _pool():
    from utlpfor_amd import workloads
    return [(workloads.plain_1m(50000, seed=21), 0), (workloads.delta_postings(50000, seed=22), 1),
            (workloads.timestamps64(20000, seed=23), 0), (workloads.uniform64(10000, seed=24), 0),
            (workloads.delta_postings(5000, seed=25), 5)]


def _oracle_store(orc, store):
    """The oracle's decode of every store block: values and status words."""
    data = store.data.to_numpy(store.nbytes)
    boff = store.block_off.to_numpy()
    ooff = store.out_off.to_numpy()
    res = {}
    for is64 in (False, True):
        res[is64] = orc.unpack_blocks(data, boff, ooff, int(ooff[-1]), is64)
    return data, boff, ooff, res


def _check_against_oracle(orc, store, oracle_state, r):
    from utlpfor_amd import serve
    data, boff, ooff, res = oracle_state
    if r.kind == serve.DECOMPRESS:
        st = store.streams[r.stream]
        vals, sts = res[st.is64]
        b0 = st.first_block + r.first
        want_st = sts[b0:b0 + r.nblocks]
        assert np.array_equal(r.status & 0xFF, want_st & 0xFF), (r.stream, r.first, r.nblocks)
        v0 = int(ooff[b0])
        want = vals[v0:v0 + r.result.size]
        for k in range(r.nblocks):
            if want_st[k] == 0:
                a, b = k * 128, min((k + 1) * 128, r.result.size)
                assert np.array_equal(r.result[a:b], want[a:b]), (r.stream, r.first, k)
    elif r.kind == serve.GET:
        st = store.streams[r.stream]
        get = orc.GetUint64 if st.is64 else orc.GetUint32
        for p, v, s in zip(r.positions, r.result, r.status):
            b = st.first_block + int(p) // 128
            wv, we = get(int(p) % 128, data[int(boff[b]):int(boff[b + 1])].tobytes())
            assert (s & 0xFF) == (1 if we == 99 else we), (r.stream, int(p))
            if we == 0:
                assert v == wv
    else:
        want, _, _ = orc.pack_stream(r.flag, r.values)
        assert np.array_equal(r.result, want)


def test_serving_mix_equals_alone_and_oracle(ctx, orc):
    from utlpfor_amd import serve
    store = _store(ctx)
    say(f"store: {len(store.streams)} streams, {store.nblocks} blocks, {store.nbytes} bytes")
    reqs = serve.make_requests(store, _pool(), n=240, rate=2e5, seed=7, min_values=1, max_values=20000)
    eng = serve.Engine(ctx, store, heap_bytes=64 << 20)
    stats = eng.serve(reqs)
    say(f"mixed: {stats}")
    assert stats["mean_batch"] > 2  # requests really shared ticks
    mixed = [(r.result.copy(), None if r.status is None else r.status.copy()) for r in reqs]
    for r, (res, sts) in zip(reqs, mixed):
        eng.run_batch([r])
        assert np.array_equal(r.result, res), (r.kind, r.stream)
        if sts is not None:
            assert np.array_equal(r.status, sts)
    state = _oracle_store(orc, store)
    for r in reqs:
        _check_against_oracle(orc, store, state, r)
    kinds = {r.kind for r in reqs}
    assert kinds == {serve.DECOMPRESS, serve.COMPRESS, serve.GET}
    store.free()


def test_serving_admission_by_heap(ctx):
    """A heap that holds only a few requests splits the work into more ticks
    and still answers every request correctly; the same heap allowed to grow
    with the device's free memory takes fewer ticks, with the same results."""
    from utlpfor_amd import serve
    store = _store(ctx)
    reqs = serve.make_requests(store, _pool(), n=60, rate=1e6, seed=9, min_values=1000, max_values=20000)
    big = serve.Engine(ctx, store, heap_bytes=64 << 20)
    a = big.serve(reqs)
    want = [r.result.copy() for r in reqs]
    small_heap = max(big.cost(r) for r in reqs) * 2
    small = serve.Engine(ctx, store, heap_bytes=small_heap, grow=False)
    b = small.serve(reqs)
    say(f"big heap: {a['ticks']} ticks, small heap ({small_heap} B): {b['ticks']} ticks")
    assert b["ticks"] > a["ticks"]
    for r, w in zip(reqs, want):
        assert np.array_equal(r.result, w)
    # the same small heap allowed to grow with the device's free memory
    grown = serve.Engine(ctx, store, heap_bytes=small_heap, grow=True)
    c = grown.serve(reqs)
    say(f"growing heap: {c['ticks']} ticks, {c['heap_growths']} growths")
    assert c["heap_growths"] >= 1 and c["ticks"] < b["ticks"]
    for r, w in zip(reqs, want):
        assert np.array_equal(r.result, w)
    for e in (big, small, grown):
        e.close()
    store.free()


def test_host_stage_chunks_and_threads(ctx):
    """Result delivery through the pinned staging buffer: a buffer much
    smaller than the range (many chunks, destinations straddling chunk
    ends) and pieces small enough that the thread pool copies most of them.
    Every destination receives exactly its device bytes."""
    from utlpfor_amd import serve
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(71)
    n = 3_000_001
    src = rng.integers(0, 256, n, dtype=np.uint8)
    d = DeviceArray.from_numpy(src, ctx)
    stage = serve._HostStage(ctx, nbytes=100_003, threads=4)
    stage.piece = 4096
    cuts = np.sort(rng.choice(np.arange(1, n), 500, replace=False))
    edges = np.concatenate([[0], cuts, [n]])
    dsts = []
    for a, b in zip(edges[:-1], edges[1:]):
        if rng.random() < 0.8:  # leave some gaps
            dt = np.uint32 if (b - a) % 4 == 0 and rng.random() < 0.5 else np.uint8
            dsts.append((int(a), np.empty((b - a) // np.dtype(dt).itemsize, dt)))
    ctx.sync()
    stage.fetch(d.ptr, n, list(reversed(dsts)))
    for off, arr in dsts:
        assert np.array_equal(arr.view(np.uint8), src[off:off + arr.nbytes]), off
    stage.close()


def test_host_stage_put(ctx):
    """Uploads through the pinned staging buffer: arrays of several dtypes
    at their offsets, across many small chunks, land byte for byte."""
    from utlpfor_amd import serve
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(72)
    n = 2_000_003
    want = rng.integers(0, 256, n, dtype=np.uint8)
    edges = np.concatenate([[0], np.sort(rng.choice(np.arange(1, n), 300, replace=False)), [n]])
    srcs = []
    for a, b in zip(edges[:-1], edges[1:]):
        piece = want[a:b]
        srcs.append((int(a), piece.view(np.uint32).copy() if (b - a) % 4 == 0 else piece.copy()))
    d = DeviceArray(n, np.uint8, ctx)
    stage = serve._HostStage(ctx, nbytes=65_537, threads=4)
    stage.piece = 2048
    stage.put(d.ptr, n, srcs[::-1])
    ctx.sync()
    assert np.array_equal(d.to_numpy(), want)
    stage.close()
