"""Where the batched u32 decode writes: streams of arbitrary lengths (blocks
at odd value offsets, partial blocks) decoded into output buffers at every
16-byte phase, by the flat and the persistent kernel; failed blocks inside
otherwise whole rounds; and output slots wider than their blocks (decode,
and Get by global position over that index). Values
outside the decoded slots, and the slots of failed blocks, keep the bytes
the buffer held before (the C-ABI's contract: values of failed blocks are
not written)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POISON = 0xA5A5A5A5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from utlpfor_amd.device import Context
    return Context(0)


def _uneven(seed, nstreams=400, top=3000):
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0, np.log(top + 1), nstreams))).astype(np.int64)
    flags = rng.choice([0, 1, 2, 4, 3], lens.size).astype(np.uint8)
    parts = []
    for L in lens:
        if rng.random() < 0.5:
            parts.append(rng.integers(0, 1 << int(rng.integers(0, 33)), int(L), dtype=np.uint64).astype(np.uint32))
        else:
            parts.append((np.cumsum(rng.integers(-1024, 1025, int(L))) + 7000).astype(np.int64).astype(np.uint32))
    return np.concatenate(parts), lens, flags


def _decode_raw(ctx, s, out_ptr, status, persistent):
    from utlpfor_amd import _lib
    from utlpfor_amd.device import check
    lib = _lib.load()
    if persistent:
        code = lib.utl_decode_persistent(ctx.handle, 0, s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks,
                                         out_ptr, status.ptr, None)
    else:
        code = lib.utl_decode_u32(ctx.handle, s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks, out_ptr,
                                  status.ptr, None)
    check(code, "decode")


@pytest.mark.parametrize("persistent", [False, True])
def test_uneven_streams_every_output_phase(ctx, persistent):
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    v, lens, flags = _uneven(21)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    ooff = s.out_off.to_numpy()
    odd = int(np.count_nonzero(ooff[:-1] & 1))
    say(f"store paths: {lens.size} streams, {v.size} values, {s.nblocks} blocks ({odd} at odd offsets), "
        f"persistent={persistent}")
    assert odd > s.nblocks // 4
    pad = 40
    out = DeviceArray(v.size + 2 * pad, np.uint32, ctx)
    st = DeviceArray(s.nblocks, np.int32, ctx)
    for shift in (0, 1, 2, 3, 5, 13, 16, 31):
        out.fill_bytes(0xA5)
        ctx.sync()
        _decode_raw(ctx, s, out.ptr + 4 * shift, st, persistent)
        ctx.sync()
        got = out.to_numpy()
        assert np.all(st.to_numpy() == 0), shift
        assert np.array_equal(got[shift:shift + v.size], v), shift
        assert np.all(got[:shift] == POISON) and np.all(got[shift + v.size:] == POISON), shift
    s.free()


@pytest.mark.parametrize("persistent", [False, True])
def test_failed_blocks_are_not_written(ctx, orc, persistent):
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(22)
    v, lens, flags = _uneven(23, nstreams=200)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    data = s.bytes_host().copy()
    boff, ooff = s.block_off.to_numpy(), s.out_off.to_numpy()
    # a header byte of every 17th block: count > 128, bad type, bw > 32 ...
    bad = np.arange(3, s.nblocks, 17)
    for k in bad:
        data[int(boff[k]) + int(rng.integers(0, 4))] ^= int(rng.integers(1, 256))
    meta = s.meta.to_numpy()
    cs = batch.from_host_blocks(ctx, data, meta & 0xFFFFFF, meta >> 24)
    wout, wst = orc.unpack_blocks(data, boff, ooff, s.nvalues, False)
    nfail = int(np.count_nonzero(wst))
    say(f"failed blocks: {nfail} of {s.nblocks}, persistent={persistent}")
    assert nfail > 0
    out = DeviceArray(s.nvalues, np.uint32, ctx)
    out.fill_bytes(0xA5)
    st = DeviceArray(s.nblocks, np.int32, ctx)
    ctx.sync()
    _decode_raw(ctx, cs, out.ptr, st, persistent)
    ctx.sync()
    got, gst = out.to_numpy(), st.to_numpy()
    wcode = np.where((wst & 0xFF) == 99, 1, wst & 0xFF)
    assert np.array_equal(gst & 0xFF, wcode)
    # a decoded block writes its count (header byte 0) values at its slot;
    # nothing else changes
    want = np.full(s.nvalues, POISON, np.uint32)
    for k in np.nonzero(wst == 0)[0]:
        a, c = int(ooff[k]), int(data[int(boff[k])])
        want[a:a + c] = wout[a:a + c]
    assert np.array_equal(got, want)
    s.free()
    cs.free()


def test_slots_wider_than_blocks(ctx):
    """A caller's out_off may leave room after a block: the gap keeps its
    bytes and every block lands at its own slot."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(24)
    v, lens, flags = _uneven(25, nstreams=150)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    ooff = s.out_off.to_numpy()
    counts = np.diff(ooff)
    gap = np.where(rng.random(s.nblocks) < 0.1, rng.integers(1, 6, s.nblocks), 0).astype(np.uint64)
    wide = np.concatenate([[0], np.cumsum(counts + gap)]).astype(np.uint64)
    s2 = batch.DeviceStream(ctx, s.data, s.meta, s.block_off, DeviceArray.from_numpy(wide, ctx), s.nblocks,
                            int(wide[-1]), s.nbytes, False)
    out = DeviceArray(int(wide[-1]), np.uint32, ctx)
    out.fill_bytes(0xA5)
    st = DeviceArray(s.nblocks, np.int32, ctx)
    ctx.sync()
    say(f"wide slots: {s.nblocks} blocks, {int(gap.sum())} gap values")
    _decode_raw(ctx, s2, out.ptr, st, False)
    ctx.sync()
    got = out.to_numpy()
    assert np.all(st.to_numpy() == 0)
    for k in range(s.nblocks):
        a, c = int(wide[k]), int(counts[k])
        assert np.array_equal(got[a:a + c], v[int(ooff[k]):int(ooff[k]) + c]), k
        assert np.all(got[a + c:int(wide[k + 1])] == POISON), k
    # Get by global position over the same index: a position in a block's
    # slot reads its value; one in the gap after it is past that block's
    # count (PositionOutOfRange). The directory-less search starts at
    # pos / 128, which overshoots here and must walk back.
    pos = rng.integers(0, int(wide[-1]), 20000).astype(np.int64)
    blk = np.searchsorted(wide, pos, side="right") - 1
    inb = pos - wide[blk].astype(np.int64)
    valid = inb < counts[blk].astype(np.int64)
    for use_dir in (True, False):
        vals, gst = batch.get_batch(ctx, s2, DeviceArray.from_numpy(pos, ctx), use_dir=use_dir)
        ctx.sync()
        gv, gs = vals.to_numpy(), gst.to_numpy()
        assert np.all(gs[valid] == 0) and np.all(gs[~valid] == 2)  # ErrPositionOutOfRange
        assert np.array_equal(gv[valid], v[(ooff[blk] + inb)[valid].astype(np.int64)])
    s2.pos_dir.free()
    s2.out_off.free()
    s.free()
