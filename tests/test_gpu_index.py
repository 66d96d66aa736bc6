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
