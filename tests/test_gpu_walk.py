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
    starts, lens, counts = [], [], []
    pos, st = 0, 0
    while pos < len(buf):
        bl, err = orc.BlockLength(buf[pos:])
        if err == 0 and bl > len(buf) - pos:
            err = 1
        if err:
            st = err
            break
        starts.append(pos)
        lens.append(bl)
        counts.append(buf[pos])
        pos += bl
    return starts, lens, counts, st


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
