"""GPU index of foreign streams (block starts unknown) against a host serial
BlockLength walk over the oracle (block_length.go:8, example_test.go:171-187).

Every stream is walked from its first byte: block k+1 starts BlockLength(block
k) bytes after block k, and the walk stops at the stream's end, at the first
BlockLength error, or at a block longer than the bytes left (ErrInvalidBuffer).
The device walk must give the same block starts, lengths, counts and stop
statuses, and decoding through its index must equal the oracle's decode of
the same byte ranges.
"""
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
    c = Context(0)
    say(f"device: {c.device_info()}")
    return c


def host_walk(orc, buf: bytes):
    """Serial walk of one stream: (starts, lengths, counts, status)."""
    st, ln, ct, status = orc.walk_stream(np.frombuffer(buf, np.uint8))
    return [int(x) for x in st], [int(x) for x in ln], [int(x) for x in ct], status


def run_walk(ctx, orc, streams, is64):
    """Walks `streams` (list of byte strings) on the device and checks the
    index against host_walk and the decode against the oracle."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    sizes = np.array([len(b) for b in streams], dtype=np.uint64)
    stream_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    blob = np.frombuffer(b"".join(streams), dtype=np.uint8)
    data = DeviceArray.from_numpy(np.concatenate([blob, np.zeros(16, np.uint8)]), ctx)
    s, counts, status = batch.index_foreign(ctx, data, stream_off, is64)
    want_starts, want_lens, want_counts = [], [], []
    for i, b in enumerate(streams):
        st0, ln, ct, st = host_walk(orc, b)
        assert counts[i] == len(st0), (i, counts[i], len(st0))
        assert status[i] == st, (i, status[i], st)
        want_starts += [int(stream_off[i]) + p for p in st0]
        want_lens += ln
        want_counts += ct
    nb = len(want_starts)
    assert s.nblocks == nb
    boff = s.block_off.to_numpy()
    meta = s.meta.to_numpy(nb)
    assert np.array_equal(boff[:nb], np.array(want_starts, dtype=np.uint64))
    assert np.array_equal(meta & 0xFFFFFF, np.array(want_lens, dtype=np.uint32))
    assert np.array_equal(meta >> 24, np.array(want_counts, dtype=np.uint32))
    if nb:
        assert boff[nb] == want_starts[-1] + want_lens[-1]
    ooff = s.out_off.to_numpy()
    assert np.array_equal(ooff, np.concatenate([[0], np.cumsum(want_counts)]).astype(np.uint64))
    out, dst = batch.decode(ctx, s)
    ctx.sync()
    wout, wst = orc.unpack_blocks(blob, boff, ooff, int(ooff[-1]), is64)
    gst = dst.to_numpy(nb)
    wcodes = np.where((wst & 0xFF) == 99, 1, wst)
    assert np.array_equal(gst, wcodes)
    gout = out.to_numpy(int(ooff[-1]))
    ok = np.repeat(wst == 0, np.diff(ooff).astype(np.int64))
    assert np.array_equal(gout[ok], wout[ok])
    s.free()
    return counts, status


def make_streams(orc, rng, nstreams, max_len, is64, flags=(0, 1, 2, 4, 3)):
    streams, values = [], []
    for _ in range(nstreams):
        n = int(rng.integers(0, max_len))
        if is64:
            v = (int(1.7e18) + np.cumsum(rng.geometric(1 / 1000, n))).astype(np.uint64)
            if rng.random() < 0.5:
                v = rng.integers(0, 1 << 64, n, dtype=np.uint64)
        else:
            v = rng.integers(0, 1 << int(rng.integers(1, 33)), n, dtype=np.uint64).astype(np.uint32)
        f = int(rng.choice(flags))
        streams.append(orc.pack_stream(f, v)[0].tobytes() if n else b"")
        values.append(v)
    return streams, values


def test_walk_valid_streams_u32(ctx, orc):
    rng = np.random.default_rng(21)
    streams, values = make_streams(orc, rng, 120, 6000, False)
    streams.append(orc.pack_stream(1, np.cumsum(rng.integers(0, 64, 300000)).astype(np.uint32))[0].tobytes())
    say(f"walk u32: {len(streams)} streams, {sum(map(len, streams))} bytes")
    counts, status = run_walk(ctx, orc, streams, False)
    assert np.all(status == 0)


def test_walk_valid_streams_u64(ctx, orc):
    rng = np.random.default_rng(22)
    streams, _ = make_streams(orc, rng, 40, 3000, True, flags=(0, 1, 2))
    say(f"walk u64: {len(streams)} streams, {sum(map(len, streams))} bytes")
    counts, status = run_walk(ctx, orc, streams, True)
    assert np.all(status == 0)


def test_walk_truncated_and_corrupt(ctx, orc):
    rng = np.random.default_rng(23)
    streams, _ = make_streams(orc, rng, 60, 3000, False)
    cases = []
    for i, b in enumerate(streams):
        if len(b) < 40:
            continue
        kind = i % 5
        a = bytearray(b)
        if kind == 0:  # truncated last block -> ErrInvalidBuffer after the others
            a = a[:len(a) - int(rng.integers(1, 12))]
        elif kind == 1:  # count > 128 in the second block's header
            st0 = host_walk(orc, bytes(a))[0]
            if len(st0) > 1:
                a[st0[1]] = 200
        elif kind == 2:  # reserved header bit in some block
            st0 = host_walk(orc, bytes(a))[0]
            j = st0[int(rng.integers(0, len(st0)))]
            a[j + 2] |= 0x04
        elif kind == 3:  # trailing bytes shorter than a header
            a += bytes(int(rng.integers(1, 4)))
        else:  # a random byte flip anywhere
            j = int(rng.integers(0, len(a)))
            a[j] ^= 1 << int(rng.integers(0, 8))
        streams[i] = bytes(a)
        cases.append(kind)
    say(f"walk corrupt: {len(streams)} streams, kinds {sorted(set(cases))}")
    counts, status = run_walk(ctx, orc, streams, False)
    assert np.any(status != 0)


def test_walk_empty_and_tiny(ctx, orc):
    streams = [b"", b"\x01", b"", orc.PackUint32(0, np.array([7], np.uint32))[0], b"\x00\x00\x00"]
    counts, status = run_walk(ctx, orc, streams, False)
    assert list(counts) == [0, 0, 0, 1, 0]


def _starts(orc, b):
    return host_walk(orc, b)[0]


def test_walk_long_streams_chunked(ctx, orc):
    """Streams of 32 KiB and more take the chunk-parallel walk (entry votes,
    verification, re-walks of chunks whose vote was wrong). Against the
    host walk: clean long streams between short ones, a 20000-byte zero run
    between two blocks (5000 valid 4-byte count-0 blocks), a block whose
    svbLen field jumps over several chunks, a reserved header bit and a
    truncation in the middle of long streams."""
    rng = np.random.default_rng(24)
    outl = rng.integers(0, 1 << 12, 300000, dtype=np.uint64)
    k = rng.choice(outl.size, 9000, replace=False)
    outl[k] = rng.integers(1 << 20, 1 << 32, k.size, dtype=np.uint64)
    longs = [orc.pack_stream(1, np.cumsum(rng.integers(0, 64, 400000)).astype(np.uint32))[0].tobytes(),
             orc.pack_stream(0, outl.astype(np.uint32))[0].tobytes(),
             orc.pack_stream(2, rng.integers(0, 1 << 32, 100000, dtype=np.uint64).astype(np.uint32))[0].tobytes(),
             orc.pack_stream(4, rng.integers(0, 1 << 9, 250000, dtype=np.uint64).astype(np.uint32))[0].tobytes()]
    # zero run between two blocks in the middle of stream 0
    st0 = _starts(orc, longs[0])
    cut = st0[len(st0) // 2]
    zero_run = longs[0][:cut] + bytes(20000) + longs[0][cut:]
    # stream 1: a block with exceptions gets svbLen 60000
    a = bytearray(longs[1])
    st1 = _starts(orc, longs[1])
    jb = next(p for p in st1[len(st1) // 3:] if a[p + 3] != 0)
    a[jb + 4:jb + 6] = (60000).to_bytes(2, "little")
    jump = bytes(a)
    # stream 2: reserved bit in a block two thirds in
    a = bytearray(longs[2])
    st2 = _starts(orc, longs[2])
    a[st2[2 * len(st2) // 3] + 2] |= 0x04
    reserved = bytes(a)
    truncated = longs[3][:-7]
    shorts, _ = make_streams(orc, rng, 30, 3000, False)
    streams = []
    for i, b in enumerate([longs[0], zero_run, longs[1], jump, longs[2], reserved, longs[3], truncated]):
        streams += [shorts[i], b, shorts[i + 8]]
    streams += shorts[16:]
    say(f"walk long: {len(streams)} streams, {sum(map(len, streams))} bytes")
    counts, status = run_walk(ctx, orc, streams, False)
    big = [i for i, b in enumerate(streams) if len(b) >= 32768]
    assert len(big) >= 8
    assert status[streams.index(reserved)] == 4
    assert status[streams.index(truncated)] == 1
    assert counts[streams.index(zero_run)] == counts[streams.index(longs[0])] + 5000


def test_walk_many_streams_long_in_every_tile(ctx, orc):
    """More streams than one tile of the chunk-offset scan (1024 streams):
    long streams in the first, middle and last tiles, the last stream long,
    between short and empty ones, so that every tile offset is used. Two of
    them have more than 64 chunks (the wide verification), one of those a
    20000-byte zero run in the middle (5000 valid 4-byte blocks, whose
    entry votes can miss and are walked again)."""
    rng = np.random.default_rng(26)
    streams, _ = make_streams(orc, rng, 3000, 200, False)
    for i in (3, 1023, 1024, 2047):
        n = int(rng.integers(40000, 120000))
        streams[i] = orc.pack_stream(int(rng.choice([0, 1, 2])),
                                     np.cumsum(rng.integers(0, 1 << 8, n)).astype(np.uint32))[0].tobytes()
    streams[2999] = orc.pack_stream(0, rng.integers(0, 1 << 16, 600000, dtype=np.uint64).astype(np.uint32))[0].tobytes()
    big = orc.pack_stream(2, rng.integers(0, 1 << 16, 600000, dtype=np.uint64).astype(np.uint32))[0].tobytes()
    st0 = _starts(orc, big)
    cut = st0[len(st0) // 2]
    streams[1500] = big[:cut] + bytes(20000) + big[cut:]
    say(f"walk many: {len(streams)} streams, {sum(map(len, streams))} bytes")
    counts, status = run_walk(ctx, orc, streams, False)
    assert np.all(status == 0)
    assert sum(len(b) >= 32768 for b in streams) == 6
    assert min(len(streams[1500]), len(streams[2999])) > 65 * 16320
    assert counts[1500] == len(st0) + 5000


def test_walk_dense_short_streams(ctx, orc):
    """Short streams of many small blocks (zeros: 4-byte blocks; values below
    16: 68-byte blocks), which the plan hands to the workgroup walk by
    pointer jumping, clean, damaged (a middle header byte, a truncation, a
    zero run) and with a first block unlike the rest (a small block before
    large ones, a large one before small ones), between ordinary short
    streams, a long stream of 68-byte blocks and long streams of 4-byte
    blocks (chunks walked by workgroups), one of those damaged in a middle
    header: the index equals the host walk stream by stream, and so does
    the decode through it."""
    rng = np.random.default_rng(31)
    zeros = orc.pack_stream(0, np.zeros(100_000, np.uint32))[0].tobytes()
    small = orc.pack_stream(2, rng.integers(0, 16, 45_000).astype(np.uint32))[0].tobytes()
    big_after_small = (orc.pack_stream(0, np.zeros(128, np.uint32))[0].tobytes()
                       + orc.pack_stream(0, rng.integers(0, 1 << 30, 800).astype(np.uint32))[0].tobytes())
    small_after_big = (orc.pack_stream(0, rng.integers(0, 1 << 30, 128).astype(np.uint32))[0].tobytes()
                       + orc.pack_stream(0, np.zeros(60_000, np.uint32))[0].tobytes())
    long_small = orc.pack_stream(2, rng.integers(0, 16, 300_000).astype(np.uint32))[0].tobytes()
    long_zeros = orc.pack_stream(0, np.zeros(1_500_000, np.uint32))[0].tobytes()
    assert len(zeros) < 32640 and len(small) < 32640 and len(long_small) > 32640 and len(long_zeros) > 32640
    long_hit = bytearray(long_zeros)
    long_hit[_starts(orc, long_zeros)[7000]] = 200  # ErrInvalidBlockLength in the second chunk
    hit = bytearray(zeros)
    st = _starts(orc, zeros)
    hit[st[len(st) // 2]] = 200  # count byte 200: ErrInvalidBlockLength
    cut = small[: len(small) - 37]
    run = small[:6800] + bytes(4000) + small[6800:]
    shorts, _ = make_streams(orc, rng, 30, 3000, False)
    streams = [zeros, small, bytes(hit), cut, run, big_after_small, small_after_big, long_small, long_zeros,
               bytes(long_hit)] + shorts
    order = rng.permutation(len(streams))
    streams = [streams[k] for k in order]
    say(f"walk dense: {len(streams)} streams, {sum(map(len, streams))} bytes")
    run_walk(ctx, orc, streams, False)


def test_walk_large_input_longer_chunks(ctx, orc):
    """A walk over more than 113 MB cuts long streams into chunks longer than
    16,320 bytes (the chunk length grows with the square root of the bytes
    walked). 4000 uneven streams encoded on the device (mixed flags, lengths
    up to 200,000 values), with the count byte of a middle block of five
    long streams set to 200: the other streams' walked blocks equal the
    encoder's index, the damaged ones stop where the oracle's walk stops."""
    from utlpfor_amd import batch, workloads
    from utlpfor_amd.device import DeviceArray
    v, lens, flags = workloads.uneven_streams(4000, seed=7)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    assert s.nbytes > 120_000_000
    data = s.bytes_host().copy()
    boff, meta = s.block_off.to_numpy(), s.meta.to_numpy()
    sb = batch.plan_blocks(lens).astype(np.int64)
    stream_off = boff[sb]
    big = np.nonzero(np.diff(stream_off) > 200_000)[0]  # more than ten chunks each
    damaged = [int(i) for i in big[::max(1, big.size // 5)][:5]]
    assert len(damaged) == 5
    for i in damaged:
        data[int(boff[sb[i] + (sb[i + 1] - sb[i]) // 2])] = 200
    say(f"walk large: {lens.size} streams, {s.nbytes} bytes, {s.nblocks} blocks, damaged {damaged}")
    d = DeviceArray.from_numpy(np.concatenate([data, np.zeros(16, np.uint8)]), ctx)
    w, counts, status = batch.index_foreign(ctx, d, stream_off)
    want_n = np.diff(sb)
    want_st = np.zeros(lens.size, np.int32)
    for i in damaged:
        st0, _, _, st = host_walk(orc, bytes(data[int(stream_off[i]):int(stream_off[i + 1])]))
        want_n[i], want_st[i] = len(st0), st
    assert np.array_equal(counts, want_n) and np.array_equal(status, want_st)
    assert all(want_st[i] == 3 for i in damaged)  # ErrInvalidBlockLength
    keep = np.concatenate([np.arange(sb[i], sb[i] + want_n[i]) for i in range(lens.size)])
    nb = keep.size
    assert w.nblocks == nb
    assert np.array_equal(w.block_off.to_numpy(nb), boff[keep])
    assert np.array_equal(w.meta.to_numpy(nb), meta[keep])
    w.free()
    d.free()
    s.free()


def test_walk_many_streams_short_ones_beside(ctx, orc):
    """A call of at least 1024 streams whose mean length is at least 32,640
    bytes walks its short streams on the context's second stream beside the
    long streams' chunk pipeline: 1500 streams (long ones of 12,000-40,000
    values, short ones, dense short ones of zeros), mixed flags, encoded on
    the device, with a middle block of a few long and short streams
    damaged; the counts and statuses equal the oracle's walk and the other
    streams' blocks equal the encoder's index."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(33)
    ns = 1500
    kind = rng.integers(0, 10, ns)
    lens = np.where(kind < 6, rng.integers(12_000, 40_000, ns),
                    np.where(kind < 9, rng.integers(1, 3000, ns), rng.integers(40_000, 200_000, ns))).astype(np.int64)
    parts = []
    for i in range(ns):
        if kind[i] == 9:  # dense short streams: zeros, 4-byte blocks
            parts.append(np.zeros(int(lens[i]), np.uint32))
        else:
            parts.append(rng.integers(0, 1 << int(rng.integers(4, 33)), int(lens[i]), dtype=np.uint64).astype(np.uint32))
    flags = rng.choice([0, 1, 2, 4, 3], ns).astype(np.uint8)
    s = batch.encode(ctx, np.concatenate(parts), stream_lengths=lens, stream_flags=flags)
    data = s.bytes_host().copy()
    boff, meta = s.block_off.to_numpy(), s.meta.to_numpy()
    sb = batch.plan_blocks(lens).astype(np.int64)
    stream_off = boff[sb]
    sizes = np.diff(stream_off)
    assert ns >= 1024 and stream_off[-1] // ns >= 32640, (stream_off[-1] // ns)
    big = np.nonzero(sizes > 70_000)[0]
    small = np.nonzero((sizes < 20_000) & (np.diff(sb) > 4))[0]
    damaged = [int(i) for i in big[:4]] + [int(i) for i in small[:4]]
    for i in damaged:
        data[int(boff[sb[i] + (sb[i + 1] - sb[i]) // 2])] = 200
    say(f"walk beside: {ns} streams, {s.nbytes} bytes, {s.nblocks} blocks, damaged {damaged}")
    d = DeviceArray.from_numpy(np.concatenate([data, np.zeros(16, np.uint8)]), ctx)
    w, counts, status = batch.index_foreign(ctx, d, stream_off)
    want_n = np.diff(sb)
    want_st = np.zeros(ns, np.int32)
    for i in damaged:
        st0, _, _, st = host_walk(orc, bytes(data[int(stream_off[i]):int(stream_off[i + 1])]))
        want_n[i], want_st[i] = len(st0), st
    assert np.array_equal(counts, want_n) and np.array_equal(status, want_st)
    keep = np.concatenate([np.arange(sb[i], sb[i] + want_n[i]) for i in range(ns)])
    nb = keep.size
    assert w.nblocks == nb
    assert np.array_equal(w.block_off.to_numpy(nb), boff[keep])
    assert np.array_equal(w.meta.to_numpy(nb), meta[keep])
    w.free()
    d.free()
    s.free()


def test_walk_long_streams_u64(ctx, orc):
    rng = np.random.default_rng(25)
    ts = (int(1.7e18) + np.cumsum(rng.geometric(1 / 1000, 200000))).astype(np.uint64)
    uni = rng.integers(0, 1 << 64, 60000, dtype=np.uint64)
    streams = [orc.pack_stream(0, ts)[0].tobytes(), orc.pack_stream(1, ts)[0].tobytes(),
               orc.pack_stream(0, uni)[0].tobytes()]
    say(f"walk long u64: {len(streams)} streams, {sum(map(len, streams))} bytes")
    counts, status = run_walk(ctx, orc, streams, True)
    assert np.all(status == 0)


@pytest.mark.parametrize("seed", range(4))
def test_walk_fuzz_damaged(ctx, orc, seed):
    """Random long and short streams (the chunk-parallel and the serial
    walk), each damaged or not at random: header bytes of random blocks
    overwritten, random byte flips, truncations, zero runs. The device
    index equals the host walk stream by stream (starts, lengths, counts,
    stop status), and so does the decode through it."""
    rng = np.random.default_rng(900 + seed)
    streams = []
    for _ in range(int(rng.integers(3, 7))):
        n = int(rng.integers(20000, 200000))
        f = int(rng.choice([0, 1, 2, 4, 3, 5]))
        if rng.random() < 0.5:
            v = np.cumsum(rng.integers(0, 1 << int(rng.integers(2, 12)), n)).astype(np.uint32)
        else:
            v = rng.integers(0, 1 << int(rng.integers(1, 33)), n, dtype=np.uint64).astype(np.uint32)
        streams.append(orc.pack_stream(f, v)[0].tobytes())
    shorts, _ = make_streams(orc, rng, 20, 3000, False)
    streams += shorts
    for i in range(len(streams)):
        a = bytearray(streams[i])
        if len(a) < 16 or rng.random() < 0.3:
            continue
        starts = _starts(orc, bytes(a))
        for _ in range(int(rng.integers(1, 4))):
            kind = int(rng.integers(0, 4))
            if kind == 0 and starts:  # a header byte of a random block
                j = starts[int(rng.integers(0, len(starts)))] + int(rng.integers(0, 4))
                if j < len(a):
                    a[j] = int(rng.integers(0, 256))
            elif kind == 1:  # a random byte flip
                j = int(rng.integers(0, len(a)))
                a[j] ^= 1 << int(rng.integers(0, 8))
            elif kind == 2:  # truncation
                a = a[:int(rng.integers(len(a) // 2, len(a)))]
            elif starts:  # a zero run before a random block
                j = starts[int(rng.integers(0, len(starts)))]
                a = a[:j] + bytes(4 * int(rng.integers(1, 3000))) + a[j:]
        streams[i] = bytes(a)
    order = rng.permutation(len(streams))
    streams = [streams[k] for k in order]
    say(f"walk fuzz seed {seed}: {len(streams)} streams, {sum(map(len, streams))} bytes, "
        f"{sum(len(b) >= 32768 for b in streams)} long")
    run_walk(ctx, orc, streams, False)
