"""Randomised differential test of the batched decode against the oracle.

Each case builds a store of streams (random flags, lengths and value
distributions, u32 or u64), damages a share of its blocks byte by byte,
and decodes random block ranges of it gathered into one index
(utl_gather_index): ranges overlap, repeat and go backwards, and the bridge
entries between them span arbitrary byte ranges, backwards too. The flat
and the persistent decode run over that index into a poisoned output.

The oracle decodes the same blocks from a copy buffer that holds each
gathered block's own bytes one after another (Unpack reads nothing outside
its block, so the bytes around it do not matter); a backward range is an
empty block there, which Unpack rejects with ErrInvalidBuffer as the
product does. Status words, overflow positions and the values of decoded
blocks must be identical, and slots of failed blocks keep the poison.
Batched Gets on the damaged stores are compared the same way, and the GPU
encoder's bytes with the oracle's on values near its decisions.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POISON = 0xA5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from utlpfor_amd.device import Context
    return Context(0)


def _values(rng, n, is64):
    k = int(rng.integers(0, 5))
    top = 64 if is64 else 32
    if k == 0:
        v = rng.integers(0, 1 << int(rng.integers(0, top)), n, dtype=np.uint64)
    elif k == 1:  # sorted with big jumps (delta, overflow near the top)
        v = np.cumsum(rng.integers(0, 1 << int(rng.integers(1, 28)), n, dtype=np.uint64))
    elif k == 2:  # small values with outliers (exceptions)
        v = rng.integers(0, 1 << 10, n, dtype=np.uint64)
        m = rng.random(n) < rng.uniform(0.01, 0.4)
        v[m] = rng.integers(1 << 20, 1 << (top - 1), int(m.sum()), dtype=np.uint64)
    elif k == 3:  # random walk (zigzag under delta)
        v = (np.cumsum(rng.integers(-5000, 5001, n)) + (1 << 20)).astype(np.int64).astype(np.uint64)
    else:
        v = rng.integers(0, 1 << 63, n, dtype=np.uint64) * 2 + rng.integers(0, 2, n, dtype=np.uint64)
    return v if is64 else (v & 0xFFFFFFFF).astype(np.uint32)


def _damaged_store(ctx, rng, is64):
    """Streams of random flags, lengths and values, encoded on the GPU, with
    a random share of the blocks damaged byte by byte; returns the device
    stream over the damaged bytes, those bytes, and the damaged share."""
    from utlpfor_amd import batch
    lens = np.floor(np.exp(rng.uniform(0, np.log(6000), int(rng.integers(20, 60))))).astype(np.int64)
    flags = rng.choice([0, 1, 2, 3, 4, 5, 6, 7], lens.size).astype(np.uint8)
    v = np.concatenate([_values(rng, int(L), is64) for L in lens])
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    data = s.bytes_host().copy()
    boff, meta = s.block_off.to_numpy(), s.meta.to_numpy()
    # damage: a few random bytes of a random share of the blocks
    share = rng.uniform(0.0, 0.3)
    for k in np.nonzero(rng.random(s.nblocks) < share)[0]:
        a, b = int(boff[k]), int(boff[k + 1])
        for _ in range(int(rng.integers(1, 4))):
            data[int(rng.integers(a, b))] = int(rng.integers(0, 256))
    cs = batch.from_host_blocks(ctx, data, meta & 0xFFFFFF, meta >> 24, is64=is64)
    s.free()
    return cs, data, lens, share


def _case(ctx, orc, seed, is64, persistent):
    from utlpfor_amd import _lib, batch
    from utlpfor_amd.device import DeviceArray, check
    lib = _lib.load()
    rng = np.random.default_rng(seed)
    cs, data, lens, share = _damaged_store(ctx, rng, is64)
    ooff = cs.out_off.to_numpy()
    # requests: random ranges, overlapping, repeated, stepping back
    rows, u, e = [], 0, 0
    for _ in range(int(rng.integers(50, 300))):
        f = int(rng.integers(0, cs.nblocks))
        c = int(rng.integers(1, min(20, cs.nblocks - f) + 1))
        rows.append((f, c, u, e))
        u += int(ooff[f + c] - ooff[f])
        e += c + 1
    req = np.array(rows, np.uint64)
    gboff, gooff = DeviceArray(e, np.uint64, ctx), DeviceArray(e, np.uint64, ctx)
    check(lib.utl_gather_index(ctx.handle, cs.block_off.ptr, cs.out_off.ptr, cs.nblocks,
                               DeviceArray.from_numpy(req.reshape(-1), ctx).ptr, len(rows), e, gboff.ptr,
                               gooff.ptr), "utl_gather_index")
    ng = e - 1
    item = 8 if is64 else 4
    out = DeviceArray(max(u, 1) * item, np.uint8, ctx)
    out.fill_bytes(POISON)
    st = DeviceArray(ng, np.int32, ctx)
    ctx.sync()
    say(f"fuzz seed {seed} {'u64' if is64 else 'u32'} persistent={persistent}: {lens.size} streams, "
        f"{cs.nblocks} blocks ({share:.0%} damaged), {len(rows)} requests, {ng} gathered blocks")
    if persistent:
        code = lib.utl_decode_persistent(ctx.handle, int(is64), cs.data.ptr, gboff.ptr, gooff.ptr, ng, out.ptr,
                                         st.ptr, None)
    else:
        f = lib.utl_decode_u64 if is64 else lib.utl_decode_u32
        code = f(ctx.handle, cs.data.ptr, gboff.ptr, gooff.ptr, ng, out.ptr, st.ptr, None)
    check(code, "decode")
    ctx.sync()
    gb, go = gboff.to_numpy(), gooff.to_numpy()
    # the oracle's copy buffer: each gathered block's own bytes
    lens_g = np.where(gb[1:] >= gb[:-1], gb[1:] - gb[:-1], 0).astype(np.int64)
    coff = np.concatenate([[0], np.cumsum(lens_g)]).astype(np.uint64)
    cbuf = np.empty(int(coff[-1]) + 16, np.uint8)
    for k in range(ng):
        a, n = int(gb[k]), int(lens_g[k])
        cbuf[int(coff[k]):int(coff[k]) + n] = data[a:a + n]
    wout, wst = orc.unpack_blocks(cbuf[:int(coff[-1])], coff, go, u, is64)
    gst = st.to_numpy()
    wcode = np.where((wst & 0xFF) == 99, 1, wst & 0xFF)
    assert np.array_equal(gst & 0xFF, wcode)
    assert np.array_equal(gst >> 8, np.where(wcode == 6, wst >> 8, 0))
    got = out.to_numpy()[:u * item].view(np.uint64 if is64 else np.uint32)
    dt = np.uint64 if is64 else np.uint32
    want = np.full(u, np.frombuffer(bytes([POISON] * item), dt)[0], dt)
    for k in np.nonzero(wst == 0)[0]:
        a = int(go[k])
        c = int(cbuf[int(coff[k])]) if lens_g[k] else 0  # count: header byte 0
        want[a:a + c] = wout[a:a + c]
    assert np.array_equal(got, want)
    cs.free()


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("is64", [False, True])
def test_fuzz_gathered_damaged(ctx, orc, seed, is64):
    _case(ctx, orc, 1000 + seed, is64, persistent=bool(seed % 2))


# every launch variant of the decode (utl_set_tune): LDS arena x waves per
# workgroup, flat and persistent (uint64 blocks have one arena size)
VARIANTS = [(False, a, w) for a in (2560, 3328, 4352) for w in (2, 4)] + [(True, 2560, w) for w in (2, 4)]


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("is64,arena,waves", VARIANTS)
def test_fuzz_every_launch_variant(ctx, orc, is64, arena, waves, persistent):
    from utlpfor_amd import batch
    before = batch.get_tune(ctx)
    _, cus = ctx.device_info()
    batch.set_tune(ctx, arena, waves, cus * 2, 3)
    say(f"launch variant: arena {arena}, {waves} waves per workgroup, persistent={persistent}")
    try:
        _case(ctx, orc, 2000 + arena + 10 * waves + int(persistent), is64, persistent)
    finally:
        batch.set_tune(ctx, before["arena_bytes"], before["waves_per_wg"], before["grid_wgs"],
                       before["tiles_per_grab"])


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("is64,arena,waves", VARIANTS)
def test_fuzz_meta_form(ctx, orc, is64, arena, waves, persistent):
    """The meta form (utl_decode_meta over the wave index) on a damaged
    store, every launch variant: statuses, overflow positions and values
    equal the oracle decoding each block's own bytes, and failed blocks
    keep the poison."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(3000 + arena + 10 * waves + int(persistent) + 7 * int(is64))
    cs, data, lens, share = _damaged_store(ctx, rng, is64)
    before = batch.get_tune(ctx)
    _, cus = ctx.device_info()
    batch.set_tune(ctx, arena, waves, cus * 2, 2)
    say(f"meta form: {'u64' if is64 else 'u32'} arena {arena}, {waves} waves, persistent={persistent}: "
        f"{cs.nblocks} blocks ({share:.0%} damaged)")
    try:
        item = 8 if is64 else 4
        out = DeviceArray(max(cs.nvalues, 1) * item, np.uint8, ctx)
        out.fill_bytes(POISON)
        st = DeviceArray(cs.nblocks, np.int32, ctx)
        wave = batch.build_wave_index(ctx, cs)
        batch.decode_meta(ctx, cs, wave, out=out, status=st, persistent=persistent)
        ctx.sync()
    finally:
        batch.set_tune(ctx, before["arena_bytes"], before["waves_per_wg"], before["grid_wgs"],
                       before["tiles_per_grab"])
    boff, ooff = cs.block_off.to_numpy(), cs.out_off.to_numpy()
    wout, wst = orc.unpack_blocks(data, boff, ooff, cs.nvalues, is64)
    gst = st.to_numpy()
    wcode = np.where((wst & 0xFF) == 99, 1, wst & 0xFF)
    assert np.array_equal(gst & 0xFF, wcode)
    assert np.array_equal(gst >> 8, np.where(wcode == 6, wst >> 8, 0))
    dt = np.uint64 if is64 else np.uint32
    got = out.to_numpy()[:cs.nvalues * item].view(dt)
    want = np.full(cs.nvalues, np.frombuffer(bytes([POISON] * item), dt)[0], dt)
    for k in np.nonzero(wst == 0)[0]:
        a = int(ooff[k])
        c = int(data[int(boff[k])])
        want[a:a + c] = wout[a:a + c]
    assert np.array_equal(got, want)
    for a in (out, st, wave):
        a.free()
    cs.free()


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("is64", [False, True])
def test_fuzz_get_damaged(ctx, orc, seed, is64):
    """Batched Get of (block, position) pairs on a damaged store, positions
    from -3 to 139 (out of range on both sides), against GetUint32 /
    GetUint64 of the oracle on each block's bytes."""
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(2000 + seed)
    cs, data, lens, share = _damaged_store(ctx, rng, is64)
    n = 4000
    blk = rng.integers(0, cs.nblocks, n).astype(np.uint64)
    pos = rng.integers(-3, 140, n).astype(np.int64)
    say(f"get fuzz seed {seed} {'u64' if is64 else 'u32'}: {cs.nblocks} blocks ({share:.0%} damaged), {n} Gets")
    vals, gst = batch.get_batch(ctx, cs, DeviceArray.from_numpy(pos, ctx), blocks=DeviceArray.from_numpy(blk, ctx))
    ctx.sync()
    gv, gs = vals.to_numpy(), gst.to_numpy()
    boff = cs.block_off.to_numpy()
    get = orc.GetUint64 if is64 else orc.GetUint32
    for i in range(n):
        k = int(blk[i])
        wv, we = get(int(pos[i]), bytes(data[int(boff[k]):int(boff[k + 1])]))
        assert (gs[i] & 0xFF) == (1 if we == 99 else we), (k, int(pos[i]), we, int(gs[i]))
        if we == 0:
            assert int(gv[i]) == wv, (k, int(pos[i]))
    cs.free()


def _edge_values(rng, n, is64):
    """Values near the encoder's decisions: exception counts around the
    list/bitmap switch (16), widths at every multiple of 4, FOR bases near
    the type's top, u64 blocks that fit 32 bits after FOR, zigzag deltas."""
    k = int(rng.integers(0, 6))
    top = 64 if is64 else 32
    if k == 0:
        bw = int(rng.integers(0, 9)) * 4
        v = rng.integers(0, 1 << bw, n, dtype=np.uint64) if bw else np.zeros(n, np.uint64)
        nexc = int(rng.integers(12, 21))
        idx = rng.choice(n, min(nexc, n), replace=False)
        v[idx] = rng.integers(1 << min(bw + 1, top - 2), 1 << (top - 1), idx.size, dtype=np.uint64)
    elif k == 1:
        base = (1 << (top - 1)) + int(rng.integers(0, 1 << 20))
        v = base + rng.integers(0, 1 << int(rng.integers(1, 24)), n, dtype=np.uint64)
    elif k == 2:
        v = (np.cumsum(rng.integers(-(1 << 12), 1 << 12, n)) + (1 << 30)).astype(np.int64).astype(np.uint64)
    elif k == 3:
        v = rng.integers(0, 1 << 32, n, dtype=np.uint64) + (np.uint64(1 << 40) if is64 else np.uint64(0))
    elif k == 4:
        v = np.full(n, rng.integers(0, 1 << (top - 1), dtype=np.uint64), np.uint64)
    else:
        v = rng.integers(0, 1 << (top - 1), n, dtype=np.uint64) * 2 + 1
    return v if is64 else (v & 0xFFFFFFFF).astype(np.uint32)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("is64", [False, True])
@pytest.mark.parametrize("method", ["slot", "direct", "sized"])
def test_fuzz_encode_bytes(ctx, orc, seed, is64, method):
    """GPU encode of many short streams with every flag combination, in every
    encode form (uint64: the slot and sized forms): the packed bytes equal the
    oracle's PackUint32/64(flag|Append) stream by stream, and the meta words
    give its block lengths and counts."""
    from utlpfor_amd import batch
    rng = np.random.default_rng(3000 + seed)
    lens = rng.integers(1, 700, int(rng.integers(30, 80)))
    flags = rng.choice(8, lens.size).astype(np.uint8)
    parts = [_edge_values(rng, int(L), is64) for L in lens]
    say(f"encode fuzz seed {seed} {'u64' if is64 else 'u32'} {method}: {lens.size} streams")
    s = batch.encode(ctx, np.concatenate(parts), stream_lengths=lens.astype(np.uint64), stream_flags=flags,
                     method=method)
    got, meta = s.bytes_host(), s.meta.to_numpy()
    off = b = 0
    for p, f in zip(parts, flags):
        want, bl, bc = orc.pack_stream(int(f), p)
        assert np.array_equal(got[off:off + want.size], want), (int(f), p.size)
        assert np.array_equal(meta[b:b + bl.size] & 0xFFFFFF, bl) and np.array_equal(meta[b:b + bl.size] >> 24, bc)
        off += want.size
        b += bl.size
    assert off == s.nbytes and b == s.nblocks
    s.free()


@pytest.mark.parametrize("is64,method", [(False, "slot"), (False, "direct"), (False, "sized"), (True, "slot"),
                                         (True, "sized")])
def test_encode_exception_heavy_blocks(ctx, orc, is64, method):
    """Blocks whose exception sections are the longest the cost model picks:
    small values with up to 100 huge ones (4-byte StreamVByte codes, about
    400 section bytes), 1- and 2-byte codes, the list/bitmap switch, at every
    byte alignment of the section (short streams of random lengths between
    them); bytes equal the oracle's stream by stream."""
    from utlpfor_amd import batch
    rng = np.random.default_rng(77 + is64)
    parts, flags = [], []
    for k in range(240):
        if k % 3 == 2:  # a short stream that shifts the next block's alignment
            parts.append(rng.integers(0, 1 << int(rng.integers(1, 20)), int(rng.integers(1, 9)), dtype=np.uint64))
        else:
            n = 128 if rng.random() < 0.8 else int(rng.integers(100, 129))
            v = rng.integers(0, 16, n, dtype=np.uint64)
            nexc = int(rng.choice([1, 15, 16, 17, 33, 64, 90, 100]))
            idx = rng.choice(n, min(nexc, n), replace=False)
            hi = int(rng.choice([12, 20, 28, 32]))
            v[idx] = rng.integers(1 << (hi - 1), 1 << hi, idx.size, dtype=np.uint64)
            parts.append(v)
        flags.append(int(rng.choice([0, 2])))  # FOR allowed or NoFOR; no delta, so outliers stay outliers
    if not is64:
        parts = [p.astype(np.uint32) for p in parts]
    lens = np.array([p.size for p in parts], np.uint64)
    say(f"exception-heavy encode {'u64' if is64 else 'u32'} {method}: {lens.size} streams")
    s = batch.encode(ctx, np.concatenate(parts), stream_lengths=lens, stream_flags=np.array(flags, np.uint8),
                     method=method)
    got = s.bytes_host()
    off = 0
    for p, f in zip(parts, flags):
        want = orc.pack_stream(int(f), p)[0]
        assert np.array_equal(got[off:off + want.size], want), (int(f), p.size)
        off += want.size
    assert off == s.nbytes
    s.free()


def test_encode_forms_with_invalid_jobs(ctx, orc):
    """A job table with counts outside 1..128 (ErrInvalidBuffer: meta word
    0xFFFFFFFF, nothing written) between valid jobs: the slot and sized
    uint32 forms give the same meta words, and every valid block's bytes at
    its scanned offset in the sized form's stream equal the oracle's
    PackUint32 of its values (the slot form's compaction copies whole tiles,
    so the blocks after a rejected one in its tile are not placed)."""
    from utlpfor_amd import _lib, batch
    from utlpfor_amd.device import DeviceArray, check
    lib = _lib.load()
    rng = np.random.default_rng(4242)
    counts = rng.integers(1, 129, 300)
    counts[[3, 77, 150, 299]] = [0, 129, 200, 255]
    flags = rng.choice(8, counts.size).astype(np.uint32)
    vals = _edge_values(rng, int(counts.sum()) + 300, False)
    in_off = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
    in_meta = (counts.astype(np.uint32) & 0xFF) | (flags << 8)
    nb = counts.size
    dv, doff, dmeta = (DeviceArray.from_numpy(a, ctx) for a in (vals, in_off, in_meta))
    metas, streams = {}, {}
    for form in ("slot", "sized"):
        meta = DeviceArray(nb, np.uint32, ctx)
        if form == "slot":
            slots = DeviceArray(nb * batch.SLOT32, np.uint8, ctx)
            check(lib.utl_encode_blocks(ctx.handle, 0, dv.ptr, doff.ptr, dmeta.ptr, nb, slots.ptr, meta.ptr), "enc")
        else:
            plan = DeviceArray(nb, np.uint64, ctx)
            check(lib.utl_encode_u32_size(ctx.handle, dv.ptr, doff.ptr, dmeta.ptr, nb, meta.ptr, plan.ptr), "size")
        boff, ooff = batch.build_index(ctx, meta, nb)
        total = int(boff.read(nb, 1)[0])
        dst = DeviceArray(total + 64, np.uint8, ctx)
        if form == "slot":
            check(lib.utl_compact(ctx.handle, 0, slots.ptr, boff.ptr, nb, dst.ptr), "compact")
            slots.free()
        else:
            check(lib.utl_encode_u32_write(ctx.handle, dv.ptr, doff.ptr, dmeta.ptr, nb, plan.ptr, boff.ptr, dst.ptr,
                                           dst.nbytes), "write")
            plan.free()
        ctx.sync()
        metas[form], bo, data = meta.to_numpy(), boff.to_numpy(), dst.to_numpy()
        streams[form] = []
        for k in range(nb):
            if 1 <= counts[k] <= 128:
                streams[form].append(data[int(bo[k]):int(bo[k + 1])].copy())
        for a in (meta, boff, ooff, dst):
            a.free()
    assert np.array_equal(metas["slot"], metas["sized"])
    bad = (counts < 1) | (counts > 128)
    assert np.all(metas["sized"][bad] == 0xFFFFFFFF)
    i = 0
    for k in range(nb):
        if bad[k]:
            continue
        v = vals[int(in_off[k]):int(in_off[k]) + int(counts[k])]
        want = np.frombuffer(bytes(orc.PackUint32(int(flags[k]) | orc.NoInPlace, v.copy())[0]), np.uint8)
        assert np.array_equal(streams["sized"][i], want), k
        i += 1
    for a in (dv, doff, dmeta):
        a.free()
