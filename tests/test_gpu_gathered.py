"""Gathered decode (utl_gather_index + one batched decode): many block ranges
of a resident stream, in any order, overlapping, repeated and going
backwards in the store, decoded in one launch. Between two requests the
gathered index holds a bridge block whose byte range may run backwards, so
a group of eight blocks can start low, jump back and continue; every block
must still be decoded from its own bytes (a round stages up to the furthest
extent of its blocks). Compared with the source values, u32 and u64."""
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


def _requests(rng, nblocks, n):
    """(first block, block count) pairs: short ranges, many overlapping and
    in random order, some repeated back to back, some stepping backwards by
    one block."""
    out = []
    while len(out) < n:
        f = int(rng.integers(0, nblocks - 1))
        c = int(rng.integers(1, min(12, nblocks - f) + 1))
        out.append((f, c))
        r = rng.random()
        if r < 0.3 and f + 1 < nblocks:
            out.append((f + 1, 1))  # starts inside the previous range
        elif r < 0.4:
            out.append((f, c))  # the same range again
        elif r < 0.5 and f > 0:
            out.append((f - 1, min(2, nblocks - f + 1)))  # one block back
    return out[:n]


@pytest.mark.parametrize("is64", [False, True])
def test_gathered_decode_backward_and_overlapping(ctx, is64):
    from utlpfor_amd import _lib, batch
    from utlpfor_amd.device import DeviceArray, check
    lib = _lib.load()
    rng = np.random.default_rng(61 + is64)
    n = 128 * 600
    if is64:
        # two-block and FOR64 blocks of several sizes
        v = rng.integers(0, 1 << 64, n, dtype=np.uint64)
        v[n // 2:] = (1 << 60) + np.cumsum(rng.integers(0, 1 << int(rng.integers(4, 30)), n - n // 2))
        dt = np.uint64
    else:
        # bit widths from 4 to 32 by block, some exceptions
        v = np.concatenate([rng.integers(0, 1 << int(b), 128, dtype=np.uint64)
                            for b in rng.integers(2, 33, n // 128)]).astype(np.uint32)
        dt = np.uint32
    s = batch.encode(ctx, v, flag=0)
    ooff = s.out_off.to_numpy()
    reqs = _requests(rng, s.nblocks, 400)
    rows = np.empty((len(reqs), 4), np.uint64)
    u = e = 0
    spans = []
    for i, (f, c) in enumerate(reqs):
        nv = int(ooff[f + c] - ooff[f])
        rows[i] = (f, c, u, e)
        spans.append((u, nv, e, c, f))
        u += nv
        e += c + 1
    d_req = DeviceArray.from_numpy(rows.reshape(-1), ctx)
    gboff = DeviceArray(e, np.uint64, ctx)
    gooff = DeviceArray(e, np.uint64, ctx)
    out = DeviceArray(max(u, 1), dt, ctx)
    st = DeviceArray(e - 1, np.int32, ctx)
    say(f"gathered {'u64' if is64 else 'u32'}: {len(reqs)} requests, {e - 1} gathered blocks, {u} values")
    check(lib.utl_gather_index(ctx.handle, s.block_off.ptr, s.out_off.ptr, s.nblocks, d_req.ptr, len(reqs), e,
                               gboff.ptr, gooff.ptr), "utl_gather_index")
    f = lib.utl_decode_u64 if is64 else lib.utl_decode_u32
    check(f(ctx.handle, s.data.ptr, gboff.ptr, gooff.ptr, e - 1, out.ptr, st.ptr, None), "gathered decode")
    ctx.sync()
    got, gst = out.to_numpy(), st.to_numpy()
    gb = gboff.to_numpy()
    backward = int(np.count_nonzero(gb[1:] < gb[:-1]))
    assert backward > 20  # the index really jumps back between requests
    for u0, nv, e0, c, f0 in spans:
        assert np.all(gst[e0:e0 + c] == 0), (f0, c)
        assert np.array_equal(got[u0:u0 + nv], v[int(ooff[f0]):int(ooff[f0]) + nv]), (f0, c)
    s.free()
