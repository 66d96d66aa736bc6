"""Device index scan (utl_index_build) against a host cumulative sum: block
byte offsets and value offsets from per-block meta words (length | count
<< 24), at sizes around every tile boundary, above the 1024-tile point where
the group-scan pass is added (several groups of 1024 tiles, the last one
partial), and from unaligned pointers (the per-item fallback kernels)."""
import ctypes as C
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


def _want(meta):
    lens = (meta & 0xFFFFFF).astype(np.uint64)
    cnts = (meta >> 24).astype(np.uint64)
    return (np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64),
            np.concatenate([[0], np.cumsum(cnts)]).astype(np.uint64))


def _meta(rng, n):
    lens = rng.integers(4, 1019, n, dtype=np.uint32)
    cnts = rng.integers(1, 129, n, dtype=np.uint32)
    return (lens | (cnts << 24)).astype(np.uint32)


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1023, 1024, 1025, 4097, 1024 * 1024, 1024 * 1024 + 7, 1025 * 1024,
                               3 * 1024 * 1024 + 5, 26 * 1024 * 1024 + 333])
def test_index_build_sizes(ctx, n):
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(n + 1)
    meta = _meta(rng, n)
    say(f"index scan: {n} blocks")
    dm = DeviceArray.from_numpy(meta if n else np.zeros(1, np.uint32), ctx)
    boff, ooff = batch.build_index(ctx, dm, n)
    ctx.sync()
    wb, wo = _want(meta)
    assert np.array_equal(boff.to_numpy(), wb) and np.array_equal(ooff.to_numpy(), wo)


@pytest.mark.parametrize("n", [5, 2049, 1024 * 1024 + 3])
def test_index_build_unaligned(ctx, n):
    """meta, block_off and out_off not 16-byte aligned: the per-item kernels."""
    from utlpfor_amd import _lib
    from utlpfor_amd.device import DeviceArray, check
    lib = _lib.load()
    rng = np.random.default_rng(7 * n)
    meta = _meta(rng, n)
    dm = DeviceArray.from_numpy(np.concatenate([np.zeros(1, np.uint32), meta]), ctx)  # meta at +4 bytes
    db = DeviceArray(n + 2, np.uint64, ctx)
    do = DeviceArray(n + 2, np.uint64, ctx)
    say(f"index scan, unaligned: {n} blocks")
    check(lib.utl_index_build(ctx.handle, dm.ptr + 4, n, db.ptr + 8, do.ptr + 8), "utl_index_build")
    ctx.sync()
    wb, wo = _want(meta)
    assert np.array_equal(db.to_numpy()[1:], wb) and np.array_equal(do.to_numpy()[1:], wo)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 1023, 1024, 1025, 4097, 1024 * 1024 + 7, 3 * 1024 * 1024 + 5,
                               26 * 1024 * 1024 + 333])
def test_wave_index_build(ctx, n):
    """utl_wave_index_build: the byte and value offsets of every eighth block."""
    from utlpfor_amd import _lib
    from utlpfor_amd.device import DeviceArray, check
    rng = np.random.default_rng(3 * n + 1)
    meta = _meta(rng, n)
    say(f"wave index: {n} blocks")
    dm = DeviceArray.from_numpy(meta, ctx)
    nw = (n + 7) // 8
    dw = DeviceArray(2 * nw + 2, np.uint64, ctx)
    dw.fill_bytes(0xEE)
    check(_lib.load().utl_wave_index_build(ctx.handle, dm.ptr, n, dw.ptr), "utl_wave_index_build")
    ctx.sync()
    wb, wo = _want(meta)
    got = dw.to_numpy()
    assert np.array_equal(got[0:2 * nw:2], wb[0:n:8]) and np.array_equal(got[1:2 * nw:2], wo[0:n:8])
    assert np.all(got[2 * nw:] == np.uint64(0xEEEEEEEEEEEEEEEE))  # nothing past the last entry


def _block_corpus(orc, rng):
    """Blocks for the header-parse pass: valid uint32 blocks of every flag
    combination and uint64 blocks (FOR64, fit-32, two-block with its
    block2Len), then damaged copies: reserved header bits (18, 19, 20), bit
    widths above 32, counts above 128 and count 0. Each is paired with
    limits: the whole block, truncations (below the header, below svbLen,
    below the payload offset of a two-block block) and extra bytes after."""
    blobs = []
    for flag in range(16):
        for n in (1, 5, 64, 128):
            v = rng.integers(0, 1 << int(rng.integers(0, 33)), n, dtype=np.uint64).astype(np.uint32)
            v[rng.random(n) < 0.1] = 0xFFFFFF00
            blobs.append(orc.PackUint32(flag, v)[0])
    for v in (np.uint64(1 << 40) + np.arange(100, dtype=np.uint64),
              np.arange(77, dtype=np.uint64),
              rng.integers(0, 1 << 64, 128, dtype=np.uint64, endpoint=False)):
        for flag in (0, 1, 4):
            blobs.append(orc.PackUint64(flag, v)[0])
    out = []
    for b in blobs:
        b = bytearray(b)
        variants = [bytes(b)]
        for bit in (18, 19, 20):
            x = bytearray(b)
            x[bit // 8] |= 1 << (bit % 8)
            variants.append(bytes(x))
        x = bytearray(b)
        x[1] = (x[1] & 0xE0) | int(rng.integers(9, 32))  # bw field 9..31 -> width 36..124
        variants.append(bytes(x))
        x = bytearray(b)
        x[0] = int(rng.integers(129, 256))
        variants.append(bytes(x))
        x = bytearray(b)
        x[0] = 0
        variants.append(bytes(x))
        for x in variants:
            for lim in {len(x), 0, 3, 4, 5, 6, 9, int(rng.integers(0, len(x) + 1)), len(x) + 7}:
                out.append((x, lim))
    return out


def test_block_lengths_pass_vs_oracle(ctx, orc):
    """utl_block_lengths (the device header-parse pass) against the oracle's
    BlockLength (block_length.go:8-65) on every (block, limit) pair: status
    and meta word (length | count << 24, 0 on error)."""
    from utlpfor_amd import _lib
    from utlpfor_amd.device import DeviceArray, check
    rng = np.random.default_rng(41)
    corpus = _block_corpus(orc, rng)
    buf = bytearray(16)
    starts, limits = [], []
    for x, lim in corpus:
        buf += bytes(int(rng.integers(0, 4)))  # arbitrary alignment
        starts.append(len(buf))
        buf += x + bytes(8)
        limits.append(starts[-1] + min(lim, len(x) + 8))
    buf += bytes(16)
    n = len(corpus)
    say(f"header-parse pass: {n} (block, limit) pairs")
    d_src = DeviceArray.from_numpy(np.frombuffer(bytes(buf), np.uint8), ctx)
    d_start = DeviceArray.from_numpy(np.array(starts, np.uint64), ctx)
    d_lim = DeviceArray.from_numpy(np.array(limits, np.uint64), ctx)
    d_meta = DeviceArray(n, np.uint32, ctx)
    d_st = DeviceArray(n, np.int32, ctx)
    check(_lib.load().utl_block_lengths(ctx.handle, d_src.ptr, d_start.ptr, d_lim.ptr, n, d_meta.ptr, d_st.ptr),
          "utl_block_lengths")
    ctx.sync()
    meta, st = d_meta.to_numpy(), d_st.to_numpy()
    seen = set()
    for k in range(n):
        src = bytes(buf[starts[k]:limits[k]])
        bl, err = orc.BlockLength(src)
        assert st[k] == err, (k, src[:8].hex(), len(src))
        want = ((bl & 0xFFFFFF) | (src[0] << 24)) if err == 0 else 0
        assert meta[k] == want, (k, src[:8].hex(), len(src))
        seen.add(err)
    assert seen >= {0, 1, 3, 4}  # OK, InvalidBuffer, InvalidBlockLength, InvalidFlags all exercised


def test_wave_index_repeated_calls(ctx):
    """The wave index's single-pass scan keeps its tile ticket and status
    words in the context's scratch: calls of alternating sizes on one
    context (the scratch refilled each time, its ticket and flags left from
    the call before) give every entry exactly."""
    from utlpfor_amd import _lib
    from utlpfor_amd.device import DeviceArray, check
    rng = np.random.default_rng(99)
    for n in (5 * 1024 * 1024 + 3, 4097, 12 * 1024 * 1024 + 1, 1, 5 * 1024 * 1024 + 3):
        meta = _meta(rng, n)
        say(f"wave index again: {n} blocks")
        dm = DeviceArray.from_numpy(meta, ctx)
        nw = (n + 7) // 8
        dw = DeviceArray(2 * nw, np.uint64, ctx)
        check(_lib.load().utl_wave_index_build(ctx.handle, dm.ptr, n, dw.ptr), "utl_wave_index_build")
        ctx.sync()
        wb, wo = _want(meta)
        got = dw.to_numpy()
        assert np.array_equal(got[0::2], wb[0:n:8]) and np.array_equal(got[1::2], wo[0:n:8])
        dm.free()
        dw.free()
