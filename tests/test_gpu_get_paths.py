"""Batched Get by stream position through the position directory
(utl_pos_dir_build + utl_get_batch_dir) against the directory-less search
of utl_get_batch and against the values: many short streams (a partial
block at the end of each, so block numbers drift far from position / 128),
boundary and out-of-range positions, arbitrary nondecreasing out_off
arrays (empty blocks, gaps, a nonzero first offset), a directory shorter
than the stream, and the directory's own entries against numpy."""
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


def _streams(seed, nstreams, top, is64=False):
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0, np.log(top + 1), nstreams))).astype(np.int64)
    flags = rng.choice([0, 1, 2, 4, 3], lens.size).astype(np.uint8)
    parts = []
    for L in lens:
        if rng.random() < 0.5:
            parts.append(rng.integers(0, 1 << int(rng.integers(0, 33)), int(L), dtype=np.uint64))
        else:
            parts.append((np.cumsum(rng.integers(-1024, 1025, int(L))) + 7000).astype(np.int64).astype(np.uint64))
    v = np.concatenate(parts)
    return (v if is64 else v.astype(np.uint32)), lens, flags


def _get(ctx, lib, s, pos, d=None, ndir=None):
    """Raw C-ABI calls: with the directory d (ndir entries) or without."""
    from utlpfor_amd.device import DeviceArray, check
    dp = DeviceArray.from_numpy(pos, ctx)
    out = DeviceArray(max(pos.size, 1), np.uint64 if s.is64 else np.uint32, ctx)
    st = DeviceArray(max(pos.size, 1), np.int32, ctx)
    if d is None:
        check(lib.utl_get_batch(ctx.handle, int(s.is64), s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks,
                                None, dp.ptr, pos.size, out.ptr, st.ptr), "get")
    else:
        check(lib.utl_get_batch_dir(ctx.handle, int(s.is64), s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks,
                                    d.ptr, ndir, dp.ptr, pos.size, out.ptr, st.ptr), "get_dir")
    ctx.sync()
    r = out.to_numpy()[:pos.size], st.to_numpy()[:pos.size]
    for a in (dp, out, st):
        a.free()
    return r


def _dir(ctx, lib, boff_dev, ooff_dev, nblocks, ndir):
    from utlpfor_amd.device import DeviceArray, check
    d = DeviceArray(ndir * 4, np.uint64, ctx)
    check(lib.utl_pos_dir_build(ctx.handle, boff_dev.ptr, ooff_dev.ptr, nblocks, d.ptr, ndir), "pos_dir")
    ctx.sync()
    return d


def _want_dir(boff, ooff, ndir):
    """Directory entries as 4 words each: block_off and out_off of the last
    block b whose out_off entry is <= 128c (0 when none), b, then the byte
    lengths (16 bits each) and value counts (8 bits each) of b and b + 1 with
    their usable bits (16, 17 of the upper half), the differences taken as
    unsigned 64-bit numbers."""
    nb = ooff.size - 1
    c = np.arange(ndir, dtype=np.uint64) * 128
    b = np.maximum(np.searchsorted(ooff[:-1], c, side="right") - 1, 0).astype(np.int64)
    l0 = boff[b + 1] - boff[b]
    n0 = ooff[b + 1] - ooff[b]
    use0 = (l0 <= 0xFFFF) & (n0 <= 0xFF)
    has1 = b + 1 < nb
    b2 = np.minimum(b + 2, nb)
    l1 = np.where(has1, boff[b2] - boff[np.minimum(b + 1, nb)], 0).astype(np.uint64)
    n1 = np.where(has1, ooff[b2] - ooff[np.minimum(b + 1, nb)], 0).astype(np.uint64)
    use1 = use0 & has1 & (l1 <= 0xFFFF) & (n1 <= 0xFF)
    z = np.uint64(0)
    lens = np.where(use0, l0, z) | np.where(use1, l1 << np.uint64(16), z)
    cnts = (np.where(use0, n0 | np.uint64(1 << 16), z) | np.where(use1, (n1 << np.uint64(8)) | np.uint64(1 << 17), z))
    return np.stack([boff[b], ooff[b], b.astype(np.uint64), lens | (cnts << np.uint64(32))], axis=1).reshape(-1)


@pytest.mark.parametrize("is64", [False, True])
def test_dir_get_equals_search_and_values(ctx, is64):
    from utlpfor_amd import _lib, batch
    lib = _lib.load()
    v, lens, flags = _streams(31 + is64, 3000, 5000, is64)
    say(f"dir get: {lens.size} streams, {v.size} values, u64={is64}")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    ooff = s.out_off.to_numpy()
    total = int(ooff[-1])
    assert total == v.size
    ndir = (total + 127) // 128
    d = s.ensure_pos_dir()
    assert d.n == 4 * ndir and np.array_equal(d.to_numpy(), _want_dir(s.block_off.to_numpy(), ooff, ndir))
    rng = np.random.default_rng(5)
    edges = np.concatenate([ooff[:-1], ooff[1:] - 1, [0, total - 1, -1, -7, total, total + 1, total + 300]])
    pos = np.concatenate([rng.integers(0, total, 200000), edges.astype(np.int64)]).astype(np.int64)
    g_dir, s_dir = _get(ctx, lib, s, pos, d, ndir)
    g_srch, s_srch = _get(ctx, lib, s, pos)
    assert np.array_equal(s_dir, s_srch) and np.array_equal(g_dir, g_srch)
    ok = (pos >= 0) & (pos < total)
    # statuses: OK inside the stream (the u64 two-block FOR quirk aside),
    # PositionOutOfRange outside
    assert np.all(s_dir[~ok] == 2)
    good = ok & (s_dir == 0)
    assert good.sum() >= 0.5 * ok.sum()
    assert np.array_equal(g_dir[good], v[pos[good]])
    # the Python wrapper takes the directory for a store of many streams
    vals, gst = batch.get_batch(ctx, s, batch.DeviceArray.from_numpy(pos, ctx))
    ctx.sync()
    assert np.array_equal(vals.to_numpy(), g_dir) and np.array_equal(gst.to_numpy(), s_dir)
    s.free()


def test_dir_over_arbitrary_out_off(ctx):
    """out_off arrays a caller may pass: empty blocks (repeated entries), gaps
    after blocks, a nonzero first entry, long runs of tiny blocks; and
    directories shorter and longer than the stream. Block contents are the
    real stream's, so each search's block gives a value or a status, and
    the two searches must agree on every position."""
    from utlpfor_amd import _lib, batch
    from utlpfor_amd.device import DeviceArray
    lib = _lib.load()
    v, lens, flags = _streams(41, 400, 3000)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    counts = np.diff(s.out_off.to_numpy()).astype(np.int64)
    rng = np.random.default_rng(9)
    for trial in range(4):
        c = counts.copy()
        c[rng.random(c.size) < 0.15] = 0                       # empty blocks
        c += np.where(rng.random(c.size) < 0.1, rng.integers(1, 200, c.size), 0)  # gaps
        if trial == 3:
            c[: c.size // 2] = np.minimum(c[: c.size // 2], 2)  # runs of tiny blocks
        first = int(rng.integers(0, 300)) if trial >= 1 else 0
        ooff = np.concatenate([[first], first + np.cumsum(c)]).astype(np.uint64)
        total = int(ooff[-1])
        s2 = batch.DeviceStream(ctx, s.data, s.meta, s.block_off, DeviceArray.from_numpy(ooff, ctx), s.nblocks,
                                total, s.nbytes, False)
        pos = np.concatenate([rng.integers(-5, total + 5, 60000), ooff.astype(np.int64),
                              ooff.astype(np.int64) - 1]).astype(np.int64)
        g_srch, s_srch = _get(ctx, lib, s2, pos)
        for ndir in ((total + 127) // 128, 1, max((total + 127) // 128 // 3, 1), (total + 127) // 128 + 50):
            say(f"trial {trial}: first {first}, total {total}, ndir {ndir}")
            d = _dir(ctx, lib, s2.block_off, s2.out_off, s2.nblocks, ndir)
            assert np.array_equal(d.to_numpy(), _want_dir(s2.block_off.to_numpy(), ooff, ndir))
            g_dir, s_dir = _get(ctx, lib, s2, pos, d, ndir)
            assert np.array_equal(s_dir, s_srch) and np.array_equal(g_dir, g_srch)
            d.free()
        s2.out_off.free()
    s.free()


def test_dir_over_blocks_with_long_byte_ranges(ctx):
    """block_off ranges longer than a directory entry's 16-bit length field:
    padding after some blocks (a caller's layout with gaps), so those blocks
    and the ones before them take the search; every Get equals the
    directory-less search's and the values are the stream's."""
    from utlpfor_amd import _lib, batch
    from utlpfor_amd.device import DeviceArray
    lib = _lib.load()
    v, lens, flags = _streams(43, 300, 4000)
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    data = s.data.to_numpy(s.nbytes)
    boff = s.block_off.to_numpy()
    rng = np.random.default_rng(12)
    pad = np.where(rng.random(s.nblocks) < 0.05, rng.integers(60000, 70000, s.nblocks), 0).astype(np.uint64)
    parts = []
    for b in range(s.nblocks):
        parts.append(data[boff[b]:boff[b + 1]])
        parts.append(np.zeros(int(pad[b]), np.uint8))
    nboff = np.concatenate([[0], np.cumsum(np.diff(boff) + pad)]).astype(np.uint64)
    ndata = np.concatenate(parts)
    say(f"long block ranges: {s.nblocks} blocks, {int((pad > 0).sum())} padded, {ndata.size} bytes")
    d_data, d_boff = DeviceArray.from_numpy(ndata, ctx), DeviceArray.from_numpy(nboff, ctx)
    s2 = batch.DeviceStream(ctx, d_data, s.meta, d_boff, s.out_off, s.nblocks, s.nvalues, ndata.size, False)
    total = s.nvalues
    ndir = (total + 127) // 128
    d = _dir(ctx, lib, s2.block_off, s2.out_off, s2.nblocks, ndir)
    assert np.array_equal(d.to_numpy(), _want_dir(nboff, s.out_off.to_numpy(), ndir))
    pos = np.concatenate([rng.integers(0, total, 100000), [0, total - 1]]).astype(np.int64)
    g_dir, s_dir = _get(ctx, lib, s2, pos, d, ndir)
    g_srch, s_srch = _get(ctx, lib, s2, pos)
    assert np.array_equal(s_dir, s_srch) and np.array_equal(g_dir, g_srch)
    good = s_dir == 0
    assert good.sum() >= 0.5 * pos.size and np.array_equal(g_dir[good], v[pos[good]])
    for a in (d, d_data, d_boff):
        a.free()
    s.free()


def _fallback_blocks(orc, rng, is64, nblocks=500):
    """Delta blocks whose Gets past row 4 take the full-unpack fallback
    (widths >= 20): zigzag deltas of a wide random walk, increasing values
    with wide steps (plain delta), and both with outlier steps (exception
    lists up to 16, bitmaps above), steps near 2^31 (32-bit values that wrap,
    so zigzag), full and partial; then damaged copies: the zigzag flag
    cleared (per-lane sums of zigzag words, which wrap for the near-2^31
    steps: ErrOverflow) or a byte changed."""
    blobs = []
    for k in range(nblocks):
        n = 128 if rng.random() < 0.7 else int(rng.integers(1, 129))
        kind = k % 5
        if kind == 0:
            v = np.cumsum(rng.integers(-(1 << 20), 1 << 20, n)) + (1 << 40)
        elif kind == 4:
            v = np.cumsum(rng.integers(1 << 29, 1 << 31, n))
        else:
            steps = rng.integers(1 << 19, 1 << 21, n)
            if kind == 2:
                idx = rng.choice(n, min(n, int(rng.integers(1, 16))), replace=False)
                steps[idx] = rng.integers(1 << 26, 1 << 28, idx.size)
            elif kind == 3:
                m = rng.random(n) < 0.3
                steps[m] = rng.integers(1 << 25, 1 << 27, int(m.sum()))
            v = np.cumsum(steps)
        v = v.astype(np.int64).astype(np.uint64)
        flag = 1 if k % 2 == 0 else 3  # Delta, Delta|NoFOR
        b = orc.PackUint64(flag, v)[0] if is64 else orc.PackUint32(flag, v.astype(np.uint32))[0]
        blobs.append(np.frombuffer(bytes(b), np.uint8).copy())
    out = list(blobs)
    for b in blobs:
        h = int.from_bytes(bytes(b[:4]), "little")
        if h & (1 << 23):  # zigzag -> plain delta over the same words
            d = b.copy()
            d[2] &= 0x7F
            out.append(d)
        d = b.copy()
        d[int(rng.integers(4, d.size))] ^= int(rng.integers(1, 256))
        out.append(d)
    return out


@pytest.mark.parametrize("is64", [False, True])
def test_full_unpack_fallback(ctx, orc, is64):
    """Every position (and three out of range) of blocks whose Gets take the
    full-unpack fallback, which runs on 16-lane groups: values and statuses
    against the oracle's GetUint32 / GetUint64, and ErrOverflow's position
    against the oracle's UnpackUint32 of the block."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(88 + is64)
    blobs = _fallback_blocks(orc, rng, is64)
    data = np.concatenate(blobs)
    lens = np.array([b.size for b in blobs], np.int64)
    counts = np.array([b[0] for b in blobs], np.int64)
    cs = batch.from_host_blocks(ctx, data, lens, counts, is64=is64)
    blk, pos = [], []
    for k, b in enumerate(blobs):
        c = int(b[0])
        p = list(range(c)) + [-1, c, 200]
        blk += [k] * len(p)
        pos += p
    blk, pos = np.array(blk, np.uint64), np.array(pos, np.int64)
    say(f"fallback gets: {len(blobs)} blocks, {pos.size} Gets, u64={is64}")
    vals, gst = batch.get_batch(ctx, cs, DeviceArray.from_numpy(pos, ctx), blocks=DeviceArray.from_numpy(blk, ctx))
    ctx.sync()
    gv, gs = vals.to_numpy(), gst.to_numpy()
    get = orc.GetUint64 if is64 else orc.GetUint32
    seen = {0: 0, 6: 0}
    for i in range(pos.size):
        b = bytes(blobs[int(blk[i])])
        wv, we = get(int(pos[i]), b)
        assert (gs[i] & 0xFF) == (1 if we == 99 else we), (int(blk[i]), int(pos[i]), we, int(gs[i]))
        if we == 0:
            assert int(gv[i]) == wv, (int(blk[i]), int(pos[i]))
        elif we == 6 and not is64:
            assert (gs[i] >> 8) == orc.UnpackUint32(b)[3], (int(blk[i]), int(pos[i]))
        if we in seen and pos[i] >= 80:
            seen[we] += 1
    # the fallback ran, on valid blocks and on overflowing ones
    assert seen[0] > 1000 and seen[6] > (100 if not is64 else 20), seen
    cs.free()
