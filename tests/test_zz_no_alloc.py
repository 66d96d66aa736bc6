"""utl_index_build_noalloc (the index scan with its scratch in a fixed
device array) against utl_index_build on the same meta words, two calls in a
row, the second with a smaller input, and its argument checks."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 67_026_944  # 1 MiB of scratch: 16 B per tile of 1024 blocks and per group of 1024 tiles


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from utlpfor_amd.device import Context
    return Context(0)


def _meta(rng, n):
    lens = rng.integers(4, 1019, n, dtype=np.uint32)
    cnts = rng.integers(1, 129, n, dtype=np.uint32)
    return (lens | (cnts << 24)).astype(np.uint32)


def _index(ctx, fn, dm, n):
    from utlpfor_amd.device import DeviceArray, check
    db = DeviceArray(n + 1, np.uint64, ctx)
    do = DeviceArray(n + 1, np.uint64, ctx)
    db.fill_bytes(0x77)
    do.fill_bytes(0x77)
    check(fn(ctx.handle, dm.ptr, n, db.ptr, do.ptr), "index")
    ctx.sync()
    r = db.to_numpy(), do.to_numpy()
    db.free()
    do.free()
    return r


def test_noalloc_equals_index_build_twice(ctx):
    from utlpfor_amd import _lib
    from utlpfor_amd.device import DeviceArray
    lib = _lib.load()
    rng = np.random.default_rng(97)
    # first a large input (the group pass: 5000 tiles), then a smaller one
    # (one workgroup's inline tile prefixes), so the second call sees the
    # first one's scratch contents
    for n in (5_000_123, 3_001):
        meta = _meta(rng, n)
        say(f"noalloc index scan: {n} blocks")
        dm = DeviceArray.from_numpy(meta, ctx)
        got = _index(ctx, lib.utl_index_build_noalloc, dm, n)
        want = _index(ctx, lib.utl_index_build, dm, n)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        lens = (meta & 0xFFFFFF).astype(np.uint64)
        assert got[0][-1] == lens.sum() and got[0][1] == lens[0]
        dm.free()


def test_noalloc_rejects_what_it_cannot_do(ctx):
    from utlpfor_amd import _lib
    from utlpfor_amd.device import DeviceArray
    lib = _lib.load()
    d = DeviceArray(64, np.uint64, ctx)
    # more blocks than the fixed scratch covers: refused before any launch
    assert lib.utl_index_build_noalloc(ctx.handle, d.ptr, MAX_BLOCKS + 1, d.ptr, d.ptr) == 100
    assert b"scratch" in lib.utl_last_error()
    # no context: creating the default one would allocate
    assert lib.utl_index_build_noalloc(None, d.ptr, 4, d.ptr, d.ptr) == 100
    d.free()
