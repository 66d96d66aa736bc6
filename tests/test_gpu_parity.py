"""GPU parity: every C-ABI entry point against the oracle, bit-exact.

Per-call Pack/Unpack/Get (uint32 and uint64) on valid blocks of every flag
combination, the reference's golden blocks, and corrupt inputs (truncation
at every cut, byte flips, the fuzz corpus); batched encode, index scan,
decode (flat and persistent) and Get gather on streams. The oracle's
REF_PANIC (the Go reference panics) must come back as ErrInvalidBuffer.
"""
import os
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from utlpfor_amd import codec
    from utlpfor_amd.device import Context
    ctx = Context(0)
    say(f"device: {ctx.device_info()}")
    return codec, ctx


def code_of(err):
    return 0 if err is None else err.code


def want_code(c):
    return 1 if c == 99 else c  # the reference's panic -> ErrInvalidBuffer


def check_unpack(codec, ctx, orc, blob, is64):
    w = (orc.UnpackUint64 if is64 else orc.UnpackUint32)(blob)
    g = (codec.UnpackUint64 if is64 else codec.UnpackUint32)(blob, ctx=ctx)
    wv, wc, we, wp = w
    gv, gc, ge = g
    assert code_of(ge) == want_code(we), (blob.hex(), we, ge)
    if we == 0:
        assert gc == wc and np.array_equal(gv, wv), blob.hex()
    elif we == 6:
        assert ge.Position == wp


def check_get(codec, ctx, orc, blob, pos, is64):
    wv, we = (orc.GetUint64 if is64 else orc.GetUint32)(pos, blob)
    gv, ge = (codec.GetUint64 if is64 else codec.GetUint32)(pos, blob, ctx=ctx)
    assert code_of(ge) == want_code(we), (pos, blob.hex(), we, ge)
    if we == 0:
        assert gv == wv, (pos, blob.hex())


def rand_values(rng, n, kind):
    if kind == 0:
        return rng.integers(0, 1 << int(rng.integers(0, 33)), n, dtype=np.uint64).astype(np.uint32)
    if kind == 1:
        return (np.cumsum(rng.integers(0, 1000, n)) + int(rng.integers(0, 1 << 31))).astype(np.uint32)
    if kind == 2:
        v = rng.integers(0, 1 << 12, n, dtype=np.uint64)
        m = rng.random(n) < 0.15
        v[m] = rng.integers(1 << 20, 1 << 32, int(m.sum()), dtype=np.uint64)
        return v.astype(np.uint32)
    if kind == 3:
        return (np.cumsum(rng.integers(-1024, 1025, n)) + 5000).astype(np.int64).astype(np.uint32)
    return rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


def test_golden_vectors_gpu(gpu, orc):
    codec, ctx = gpu
    exc = list(range(128))
    for p in (5, 15, 25, 35, 45, 55, 65, 75, 85, 95):
        exc[p] = 0x10000000 + p
    cases = [("plain_128_bw8.bin", [128 + i for i in range(128)], 0), ("plain_128_bw0.bin", [0] * 128, 0),
             ("delta_128.bin", [100 * i for i in range(128)], 1), ("exceptions_128.bin", exc, 0)]
    for name, values, flag in cases:
        say(f"golden {name}")
        golden = open(os.path.join(GOLDEN, name), "rb").read()
        packed, err = codec.PackUint32(flag, np.array(values, np.uint32), ctx=ctx)
        assert err is None and packed == golden
        out, consumed, err = codec.UnpackUint32(golden, ctx=ctx)
        assert err is None and list(out) == values and consumed == len(golden)
        for pos in range(128):
            assert codec.GetUint32(pos, golden, ctx=ctx) == (values[pos], None)


def test_pack_u32_all_flags(gpu, orc):
    codec, ctx = gpu
    rng = np.random.default_rng(101)
    for i in range(400):
        n = int(rng.choice([1, 2, 3, 7, 15, 16, 17, 31, 33, 64, 100, 127, 128, 128]))
        v = rand_values(rng, n, i % 5)
        flag = i % 16
        for extra in (0, 32):
            say(f"pack_u32 i={i} n={n} flag={flag | extra}") if i % 100 == 0 else None
            want, wwork, we = orc.PackUint32(flag | extra, v)
            work = v.copy()
            got, ge = codec.PackUint32(flag | extra, work, ctx=ctx)
            assert we == 0 and ge is None
            assert got == want, (i, flag | extra, v.tolist())
            assert np.array_equal(work, wwork), (i, flag | extra)


def test_pack_u64_all_flags(gpu, orc):
    codec, ctx = gpu
    rng = np.random.default_rng(202)
    kinds = [lambda n: rng.integers(0, 1 << 63, n, dtype=np.uint64) * 2 + 1,
             lambda n: (int(1.7e18) + np.cumsum(rng.geometric(1 / 1000, n))).astype(np.uint64),
             lambda n: rng.integers(0, 1 << 20, n, dtype=np.uint64),
             lambda n: (0x1234567800000000 | (np.arange(n, dtype=np.uint64) * 7)).astype(np.uint64),
             lambda n: (0x1000000000000000 + ((np.arange(n) % 4).astype(np.uint64) << 32)
                        + np.arange(n, dtype=np.uint64) * 12345).astype(np.uint64)]
    for i in range(200):
        n = int(rng.choice([1, 3, 16, 17, 64, 127, 128]))
        v = kinds[i % len(kinds)](n)
        flag = i % 16
        want, we = orc.PackUint64(flag, v)
        got, ge = codec.PackUint64(flag, v, ctx=ctx)
        assert we == 0 and ge is None and got == want, (i, flag)


def test_pack_invalid_lengths(gpu, orc):
    codec, ctx = gpu
    for n in (0, 129):
        assert codec.PackUint32(0, np.zeros(n, np.uint32), ctx=ctx)[1] is codec.ErrInvalidBuffer
        assert codec.PackUint64(0, np.zeros(n, np.uint64), ctx=ctx)[1] is codec.ErrInvalidBuffer


def _valid_blocks(orc, rng, count):
    blocks = []
    for i in range(count):
        n = int(rng.choice([1, 5, 16, 17, 64, 100, 128, 128, 128]))
        v = rand_values(rng, n, i % 5)
        blocks.append((orc.PackUint32(i % 16, v)[0], False))
        v64 = rng.integers(0, 1 << int(rng.integers(1, 64)), n, dtype=np.uint64)
        if i % 3 == 0:
            v64 = v64 + np.uint64(1 << 40)
        blocks.append((orc.PackUint64(i % 16, v64)[0], True))
    return blocks


def test_unpack_get_valid(gpu, orc):
    codec, ctx = gpu
    rng = np.random.default_rng(303)
    for j, (blob, is64) in enumerate(_valid_blocks(orc, rng, 80)):
        if j % 40 == 0:
            say(f"unpack/get valid j={j}")
        check_unpack(codec, ctx, orc, blob, is64)
        check_unpack(codec, ctx, orc, blob, not is64)
        count = blob[0]
        for pos in sorted({0, 1, 15, 16, 63, 64, 100, count - 1, count, -1}):
            check_get(codec, ctx, orc, blob, pos, is64)


def _corrupt_targets(orc):
    fill = lambda f: np.array([f(i) for i in range(128)], dtype=np.uint64)
    rng = np.random.default_rng(7)
    few = fill(lambda i: i % 50)
    few[7], few[50] = 0xFFFF0000, 0x00FF0000
    many = fill(lambda i: i % 50)
    for j in range(48):
        many[(j * 11) % 128] = 0xFF000000 | int(rng.integers(0, 0xFFFFFF))
    dexc = fill(lambda i: 1000 + 3 * i)
    dexc[64] = 0xFFFF0000
    defs = [(0, fill(lambda i: i % 200)), (0, few), (0, many), (1, fill(lambda i: 1000 + 3 * i)),
            (1, fill(lambda i: 1000000 + 3 * i)),
            (1, fill(lambda i: 1000 + 3 * i if i % 2 == 0 else 1000 + 3 * (i - 1) - 97)), (1, dexc),
            (1, fill(lambda i: (1 << 31) + (i * 7919) % 1000003)), (3, fill(lambda i: 5 * i))]
    t32 = [orc.PackUint32(f, v.astype(np.uint32))[0] for f, v in defs]
    t64 = [orc.PackUint64(0, fill(f))[0] for f in (lambda i: i % 200, lambda i: 0x100000000 + 3 * i,
                                                    lambda i: 0x1234567800000000 | (i * 7))]
    t64.append(orc.PackUint64(1, fill(lambda i: (1 << 60) + i * i * 977))[0])
    return t32, t64


def test_corrupt_truncation(gpu, orc):
    codec, ctx = gpu
    t32, t64 = _corrupt_targets(orc)
    positions = [0, 1, 7, 15, 16, 31, 64, 100, 127]
    for ti, blob in enumerate(t32 + t64):
        say(f"truncation target {ti} len={len(blob)}")
        for cut in range(0, len(blob) + 1):
            b = blob[:cut]
            check_unpack(codec, ctx, orc, b, False)
            check_unpack(codec, ctx, orc, b, True)
            if cut % 3 == 0 or cut == len(blob):
                for pos in positions:
                    check_get(codec, ctx, orc, b, pos, False)
                    check_get(codec, ctx, orc, b, pos, True)


def test_corrupt_byte_flips(gpu, orc):
    codec, ctx = gpu
    t32, t64 = _corrupt_targets(orc)
    rng = np.random.default_rng(9)
    for ti, blob in enumerate(t32 + t64):
        say(f"byte flips target {ti}")
        idxs = list(range(min(len(blob), 24))) + list(rng.integers(0, len(blob), 24))
        for i in idxs:
            for probe in (0x00, 0xFF, int(rng.integers(0, 256))):
                b = bytearray(blob)
                b[i] = probe
                b = bytes(b)
                check_unpack(codec, ctx, orc, b, False)
                check_unpack(codec, ctx, orc, b, True)
                for pos in (0, 7, 64, 127, 130):
                    check_get(codec, ctx, orc, b, pos, False)
                    check_get(codec, ctx, orc, b, pos, True)


def test_corrupt_headers_and_corpus(gpu, orc):
    codec, ctx = gpu
    rng = np.random.default_rng(13)
    corpus = [b"", b"\x00a800", b"0A10", b"0A00\x0e\x00" + b"0" * 64 + b"1" + b"0" * 16 + b"2" + b"0" * 81,
              b"\xff\xff\xff\xff", b"\x80\xff\xff\xff\xff\xff\xff\xff"]
    for _ in range(300):
        h = int(rng.integers(0, 1 << 32))
        body = rng.integers(0, 256, int(rng.integers(0, 700)), dtype=np.uint8).tobytes()
        corpus.append(struct.pack("<I", h) + body)
        # plausible headers: valid type, small counts, delta bits
        h2 = (h & ~(3 << 13) & ~0xFF) | (int(rng.choice([1, 2, 3])) << 13) | int(rng.integers(0, 256))
        corpus.append(struct.pack("<I", h2 & ~(0x1F << 8) | (int(rng.integers(0, 10)) << 8)) + body)
    for j, b in enumerate(corpus):
        if j % 200 == 0:
            say(f"corpus j={j}")
        check_unpack(codec, ctx, orc, b, False)
        check_unpack(codec, ctx, orc, b, True)
        for pos in (0, 7, 64, 127, 128, 135, 150, 254):
            check_get(codec, ctx, orc, b, pos, False)
            check_get(codec, ctx, orc, b, pos, True)


def test_overflow_positions_gpu(gpu, orc):
    codec, ctx = gpu
    for row in range(1, 8):
        for lane in range(0, 16, 3):
            d = np.zeros(128, np.uint32)
            d[lane], d[row * 16 + lane] = 0xFFFFFFF0, 0x20
            raw = orc.PackUint32(2 | 4, d)[0]
            h = struct.unpack_from("<I", raw)[0] | (1 << 22)
            blob = struct.pack("<I", h) + raw[4:]
            out, _, err = codec.UnpackUint32(blob, ctx=ctx)
            assert err == codec.ErrOverflow(row * 16 + lane)
            check_get(codec, ctx, orc, blob, row * 16 + lane, False)


def test_batched_stream_plain_1m(gpu, orc):
    """1,048,576 values U[0,2^12) with 3% in [2^20, 2^32), flag 0."""
    from utlpfor_amd import batch
    codec, ctx = gpu
    rng = np.random.default_rng(1)
    n = 1 << 20
    v = rng.integers(0, 1 << 12, n, dtype=np.uint64)
    idx = rng.choice(n, round(0.03 * n), replace=False)
    v[idx] = rng.integers(1 << 20, 1 << 32, idx.size, dtype=np.uint64)
    v = v.astype(np.uint32)
    say("plain 1 Mi: GPU encode")
    s = batch.encode(ctx, v, 0)
    want, bl, bc = orc.pack_stream(0, v)
    got = s.bytes_host()
    assert got.size == want.size and np.array_equal(got, want)
    say("plain 1 Mi: GPU decode")
    out, st = batch.decode(ctx, s)
    ctx.sync()
    assert np.all(st.to_numpy() == 0)
    assert np.array_equal(out.to_numpy(), v)
    out2, st2 = batch.decode(ctx, s, persistent=True)
    ctx.sync()
    assert np.array_equal(out2.to_numpy(), v) and np.all(st2.to_numpy() == 0)
    say("plain 1 Mi: Get gather")
    pos = rng.integers(0, n, 100000).astype(np.int64)
    from utlpfor_amd.device import DeviceArray
    vals, gst = batch.get_batch(ctx, s, DeviceArray.from_numpy(pos, ctx))
    ctx.sync()
    assert np.all(gst.to_numpy() == 0) and np.array_equal(vals.to_numpy(), v[pos])
    s.free()


def test_batched_mixed_streams_and_corruption(gpu, orc):
    from utlpfor_amd import batch
    codec, ctx = gpu
    rng = np.random.default_rng(5)
    lens = np.unique(np.floor(np.exp(rng.uniform(0, np.log(5001), 40))).astype(np.int64))
    flags = rng.choice([0, 1, 2, 4, 3], lens.size).astype(np.uint8)
    parts = [rand_values(rng, int(L), int(rng.integers(0, 5))) for L in lens]
    v = np.concatenate(parts)
    say(f"mixed streams: {lens.size} streams, {v.size} values")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    # oracle stream by stream
    want = np.concatenate([orc.pack_stream(int(f), p)[0] for f, p in zip(flags, parts)])
    assert np.array_equal(s.bytes_host(), want)
    out, st = batch.decode(ctx, s)
    ctx.sync()
    assert np.all(st.to_numpy() == 0) and np.array_equal(out.to_numpy(), v)
    # corrupt some blocks on the host, decode on device, compare per block
    data = s.bytes_host().copy()
    boff = s.block_off.to_numpy()
    ooff = s.out_off.to_numpy()
    for k in rng.choice(s.nblocks, min(40, s.nblocks), replace=False):
        a, b = int(boff[k]), int(boff[k + 1])
        j = int(rng.integers(a, b))
        data[j] = int(rng.integers(0, 256))
    meta = s.meta.to_numpy()
    cs = batch.from_host_blocks(ctx, data, meta & 0xFFFFFF, meta >> 24)
    cons = __import__("utlpfor_amd.device", fromlist=["DeviceArray"]).DeviceArray(s.nblocks, np.int32, ctx)
    out_c, st_c = batch.decode(ctx, cs, consumed=cons)
    ctx.sync()
    wout, wst = orc.unpack_blocks(data, boff, ooff, s.nvalues, False)
    gst = st_c.to_numpy()
    gst_codes = gst & 0xFF
    wst_codes = np.where((wst & 0xFF) == 99, 1, wst & 0xFF)
    assert np.array_equal(gst_codes, wst_codes)
    assert np.array_equal(gst >> 8, np.where(wst_codes == 6, wst >> 8, 0))
    gout = out_c.to_numpy()
    for k in np.nonzero(wst == 0)[0]:
        a, b = int(ooff[k]), int(ooff[k + 1])
        assert np.array_equal(gout[a:b], wout[a:b]), k
    s.free()
    cs.free()


def test_uneven_streams_graph_autotune(gpu, orc):
    """Mixed-flag uneven streams: GPU encode bytes equal the oracle's stream
    by stream; the autotuned persistent decode and the flat decode, each
    captured with the index scan as a HIP graph, equal the eager flat decode
    on every replay (eager work and syncs between replays)."""
    from utlpfor_amd import autotune, batch, workloads
    from utlpfor_amd.device import DeviceArray
    codec, ctx = gpu
    v, lens, flags = workloads.uneven_streams(300, max_len=20000, seed=11)
    say(f"uneven: {lens.size} streams, {v.size} values")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    got = s.bytes_host()
    off = 0
    pos = 0
    for L, f in zip(lens, flags):
        want = orc.pack_stream(int(f), v[pos:pos + int(L)])[0]
        assert np.array_equal(got[off:off + want.size], want), (int(L), int(f))
        off += want.size
        pos += int(L)
    assert off == s.nbytes
    flat_out, flat_st = batch.decode(ctx, s)
    ctx.sync()
    ref = flat_out.to_numpy()
    assert np.all(flat_st.to_numpy() == 0) and np.array_equal(ref, v)
    # every launch variant decodes the mix exactly, flat and persistent
    before = batch.get_tune(ctx)
    vout = DeviceArray(v.size, np.uint32, ctx)
    vst = DeviceArray(s.nblocks, np.int32, ctx)
    for c in autotune.candidates(ctx, s)[::7]:
        batch.set_tune(ctx, c["arena_bytes"], c["waves_per_wg"], c["grid_wgs"], c["tiles_per_grab"])
        vout.fill_bytes(0xA5)
        batch.decode(ctx, s, out=vout, status=vst, persistent=c["persistent"])
        ctx.sync()
        assert np.all(vst.to_numpy() == 0) and np.array_equal(vout.to_numpy(), v), c
    vout.free()
    vst.free()
    batch.set_tune(ctx, before["arena_bytes"], before["waves_per_wg"], before["grid_wgs"], before["tiles_per_grab"])
    tune = autotune.tune_decode(ctx, s, use_cache=False, reps=2)
    assert tune["grid_wgs"] % 8 == 0 and tune["tiles_per_grab"] >= 1
    assert tune["arena_bytes"] in batch.ARENAS32 and tune["waves_per_wg"] in batch.WAVES_PER_WG
    assert tune["ms"] == min(tune["flat_ms"], tune["persistent_ms"])
    names = {(c["arena_bytes"], c["waves_per_wg"], c["persistent"]) for c in tune["candidates"]}
    assert names == {(a, w, p) for a in batch.ARENAS32 for w in batch.WAVES_PER_WG for p in (False, True)}
    out = DeviceArray(v.size, np.uint32, ctx)
    st = DeviceArray(s.nblocks, np.int32, ctx)
    for persistent, block_index in ((True, True), (False, True), (True, False), (False, False)):
        g = batch.DecodeGraph(ctx, s, out, st, persistent=persistent, block_index=block_index)
        for rep in range(3):
            out.fill_bytes(0xA5)
            ctx.sync()
            g.launch()
            ctx.sync()
            assert np.all(st.to_numpy() == 0), (persistent, block_index, rep)
            assert np.array_equal(out.to_numpy(), ref), (persistent, block_index, rep)
            # eager work between replays
            batch.decode(ctx, s, out=flat_out, status=flat_st)
            ctx.sync()
        g.close()
    s.free()


def test_graph_replays_after_context_scratch_grows(gpu, orc):
    """A decode graph owns its scan scratch: replays stay exact after the
    context's shared scratch is freed and reallocated by a much larger index
    scan on the same context (both graph forms)."""
    from utlpfor_amd import _lib, batch, workloads
    from utlpfor_amd.device import DeviceArray, check
    codec, ctx = gpu
    lib = _lib.load()
    v, lens, flags = workloads.uneven_streams(40, max_len=5000, seed=12)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    out = DeviceArray(v.size, np.uint32, ctx)
    st = DeviceArray(s.nblocks, np.int32, ctx)
    rng = np.random.default_rng(4)
    nmax = 12 << 20
    big = DeviceArray.from_numpy((rng.integers(4, 1018, nmax) | (128 << 24)).astype(np.uint32), ctx)
    bo, oo = DeviceArray(nmax + 1, np.uint64, ctx), DeviceArray(nmax + 1, np.uint64, ctx)
    for block_index, nbig in ((True, 6 << 20), (False, nmax)):
        # each scan needs more scratch than the context held when the graph was captured
        say(f"graph with block_index={block_index}, {s.nblocks} blocks, then a {nbig}-block index scan")
        g = batch.DecodeGraph(ctx, s, out, st, persistent=True, block_index=block_index)
        check(lib.utl_index_build(ctx.handle, big.ptr, nbig, bo.ptr, oo.ptr), "index")
        ctx.sync()
        for rep in range(3):
            out.fill_bytes(0xA5)
            ctx.sync()
            g.launch()
            ctx.sync()
            assert np.all(st.to_numpy() == 0) and np.array_equal(out.to_numpy(), v), (block_index, rep)
            check(lib.utl_index_build(ctx.handle, big.ptr, nbig, bo.ptr, oo.ptr), "index")
            ctx.sync()
        g.close()
    for a in (big, bo, oo, out, st):
        a.free()
    s.free()


def test_batched_u64(gpu, orc):
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    codec, ctx = gpu
    rng = np.random.default_rng(8)
    ts = (int(1.7e18) + np.cumsum(rng.geometric(1 / 1000, 40000))).astype(np.uint64)
    uni = rng.integers(0, 1 << 64, 10000, dtype=np.uint64)
    for name, v in (("timestamps", ts), ("uniform", uni)):
        say(f"u64 {name}")
        s = batch.encode(ctx, v, 0)
        want = orc.pack_stream(0, v)[0]
        assert np.array_equal(s.bytes_host(), want)
        out, st = batch.decode(ctx, s)
        ctx.sync()
        stn = st.to_numpy()
        if name == "timestamps":
            assert np.all(stn == 0) and np.array_equal(out.to_numpy(), v)
        boff, ooff = s.block_off.to_numpy(), s.out_off.to_numpy()
        wout, wst = orc.unpack_blocks(want, boff, ooff, v.size, True)
        assert np.array_equal(stn, wst)
        ok = wst == 0
        gout = out.to_numpy()
        for k in np.nonzero(ok)[0][:200]:
            a, b = int(ooff[k]), int(ooff[k + 1])
            assert np.array_equal(gout[a:b], v[a:b])
        pos = rng.integers(0, v.size, 20000).astype(np.int64)
        vals, gs = batch.get_batch(ctx, s, DeviceArray.from_numpy(pos, ctx))
        ctx.sync()
        gsn, gv = gs.to_numpy(), vals.to_numpy()
        for i in range(0, pos.size, 97):
            k = int(np.searchsorted(ooff, pos[i], side="right") - 1)
            blob = want[int(boff[k]):int(boff[k + 1])].tobytes()
            wv, we = orc.GetUint64(int(pos[i] - ooff[k]), blob)
            assert gsn[i] == we and (we or gv[i] == wv)
        s.free()


def _u64_far_second(orc, rng, pad):
    """A uint64 two-block block whose first sub-block's svbLen is inflated by
    `pad` zero bytes, so its second sub-block starts `pad` bytes later (the
    reference accepts the longer StreamVByte area: exceptions.go:32-88)."""
    lo = rng.integers(0, 1 << 10, 128, dtype=np.uint64)
    k = rng.choice(128, 12, replace=False)
    lo[k] = rng.integers(1 << 24, 1 << 32, 12, dtype=np.uint64)
    hi = rng.integers(0, 1 << 32, 128, dtype=np.uint64)
    v = (hi << np.uint64(32)) | lo
    b = bytearray(orc.PackUint64(0, v)[0])
    h = int.from_bytes(b[:4], "little")
    assert (h >> 19) & 1 and h >> 24, "needs a two-block block with exceptions in its first sub-block"
    forb = (1 << ((h >> 15) & 3)) >> 1
    b2len = int.from_bytes(b[6 + forb:8 + forb], "little")
    c1 = len(b) - b2len
    out = b[:c1] + bytes(pad) + b[c1:]
    out[4:6] = (int.from_bytes(b[4:6], "little") + pad).to_bytes(2, "little")
    return v, bytes(out)


def test_u64_second_subblock_outside_staged_extent(gpu, orc):
    """Second sub-blocks that start past the staged extent of their block
    are restaged one at a time; several in one wavefront, between ordinary
    blocks, plus ones whose second sub-block then fails."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    codec, ctx = gpu
    rng = np.random.default_rng(31)
    blobs = []
    for i in range(24):
        if i % 3 == 2:
            blobs.append(orc.PackUint64(0, rng.integers(0, 1 << 64, 128, dtype=np.uint64))[0])
            continue
        _, b = _u64_far_second(orc, rng, int(rng.integers(1400, 3000)))
        if i % 5 == 4:  # corrupt the second sub-block's count (count mismatch / errors)
            bb = bytearray(b)
            hs = int.from_bytes(bb[:4], "little")
            forb = (1 << ((hs >> 15) & 3)) >> 1
            b2 = int.from_bytes(bb[6 + forb:8 + forb], "little")
            bb[len(bb) - b2] = 77
            b = bytes(bb)
        blobs.append(b)
    for b in blobs[:6]:
        check_unpack(codec, ctx, orc, b, True)
    data = np.frombuffer(b"".join(blobs), np.uint8)
    lens = np.array([len(b) for b in blobs], np.uint64)
    boff = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    ooff = (np.arange(len(blobs) + 1) * 128).astype(np.uint64)
    s = batch.from_host_blocks(ctx, data, lens.astype(np.uint32), np.full(len(blobs), 128, np.uint32), is64=True)
    cons = DeviceArray(len(blobs), np.int32, ctx)
    out, st = batch.decode(ctx, s, consumed=cons)
    ctx.sync()
    wout, wst = orc.unpack_blocks(data, boff, ooff, 128 * len(blobs), True)
    gst = st.to_numpy()
    assert np.array_equal(gst, np.where((wst & 0xFF) == 99, 1, wst)), (gst, wst)
    assert np.any(gst == 0) and np.any(gst != 0)
    gout, gcons = out.to_numpy(), cons.to_numpy()
    for k in np.nonzero(wst == 0)[0]:
        assert np.array_equal(gout[128 * k:128 * k + 128], wout[128 * k:128 * k + 128]), k
        assert gcons[k] == lens[k], k
    s.free()
