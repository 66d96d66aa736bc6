"""Full-size parity of the batched encode and decode against the oracle.

The benchmark workloads at their stated sizes, byte for byte: the GPU
encoder's concatenated stream of the 64 Mi delta postings (flag Delta)
equals the oracle's PackUint32(Delta|Append) stream block by block, and its
meta words equal the oracle's block lengths and counts; 20,000 uneven
streams of the mixed-flag generator (326 M values) are compared stream by
stream in every encode form (the oracle packs them on a thread pool), and
all 200,000 (3.27 G values) in the default form, with the graph step's
decode of every block against the oracle's; the two uint64 workloads
(32 Mi FOR64 timestamps, 8 Mi uniform two-block values) compare encode bytes
and the decode of every block (status words and values) with the oracle's.
(format_conformance_test.go:727-754 is the reference's byte-exact check on
its golden blocks; this is the same check on the workloads.)
"""
import sys
from concurrent.futures import ThreadPoolExecutor

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


def _meta_of(bl, bc):
    return (bl.astype(np.uint32) & 0xFFFFFF) | (bc.astype(np.uint32) << 24)


@pytest.mark.parametrize("method", ["slot", "direct", "sized"])
def test_delta64m_encode_bytes_and_decode(ctx, orc, method):
    from utlpfor_amd import batch, workloads
    v = workloads.delta_postings(64 << 20)
    say(f"delta postings: {v.size} values; oracle pack")
    want, bl, bc = orc.pack_stream(orc.Delta, v, nthreads=16)
    say(f"GPU encode ({method})")
    s = batch.encode(ctx, v, flag=1, method=method)
    assert s.nbytes == want.size and s.nblocks == bl.size
    assert np.array_equal(s.meta.to_numpy(), _meta_of(bl, bc))
    assert np.array_equal(s.bytes_host(), want)
    out, st = batch.decode(ctx, s)
    ctx.sync()
    assert np.all(st.to_numpy() == 0) and np.array_equal(out.to_numpy(), v)
    out.free()
    st.free()
    s.free()


@pytest.mark.parametrize("method", ["slot", "direct", "sized"])
def test_uneven_20000_streams_encode_bytes(ctx, orc, method):
    from utlpfor_amd import batch, workloads
    v, lens, flags = workloads.uneven_streams(20_000, seed=3)
    vo = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    say(f"uneven: {lens.size} streams, {v.size} values; oracle pack")

    def pack(i):
        return orc.pack_stream(int(flags[i]), v[vo[i]:vo[i + 1]], nthreads=1)

    with ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(pack, range(lens.size)))
    say(f"GPU encode ({method})")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags, method=method)
    got = s.bytes_host()
    meta = s.meta.to_numpy()
    boff = 0
    bk = 0
    for i, (want, bl, bc) in enumerate(parts):
        n = want.size
        assert np.array_equal(got[boff:boff + n], want), (i, int(lens[i]), int(flags[i]))
        assert np.array_equal(meta[bk:bk + bl.size], _meta_of(bl, bc)), i
        boff += n
        bk += bl.size
    assert boff == s.nbytes and bk == s.nblocks
    out, st = batch.decode(ctx, s)
    ctx.sync()
    assert np.all(st.to_numpy() == 0) and np.array_equal(out.to_numpy(), v)
    out.free()
    st.free()
    s.free()


def test_uneven_200000_streams_whole_workload(ctx, orc):
    """The headline workload at its full size (200,000 streams, 3.27 G
    values, 25.7 M blocks, 7.8 GB packed): the GPU encoder's bytes equal the
    oracle's PackUint32(flag|Append) streams and its meta words their block
    lengths and counts; the GPU decode of every block (the graph step: wave
    index + meta-form decode) equals the oracle's UnpackUint32 of every
    block, values and statuses, and the generated values."""
    from utlpfor_amd import batch, workloads
    from utlpfor_amd.device import DeviceArray
    v, lens, flags = workloads.uneven_streams(200_000, seed=3)
    vo = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    say(f"uneven: {lens.size} streams, {v.size} values; oracle pack")

    def pack(i):
        return orc.pack_stream(int(flags[i]), v[vo[i]:vo[i + 1]], nthreads=1)

    with ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(pack, range(lens.size)))
    want = np.concatenate([p[0] for p in parts])
    wmeta = np.concatenate([_meta_of(p[1], p[2]) for p in parts])
    del parts
    say("GPU encode")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    assert s.nbytes == want.size and s.nblocks == wmeta.size
    assert np.array_equal(s.meta.to_numpy(), wmeta)
    got = s.bytes_host()
    assert np.array_equal(got, want)
    del want, got
    say("GPU decode (graph step)")
    out = DeviceArray(v.size, np.uint32, ctx)
    st = DeviceArray(s.nblocks, np.int32, ctx)
    g = batch.DecodeGraph(ctx, s, out, st, persistent=False, block_index=False)
    g.launch()
    ctx.sync()
    g.close()
    gout, gst = out.to_numpy(), st.to_numpy()
    out.free()
    st.free()
    say("oracle decode of every block")
    wout, wst = orc.unpack_blocks(s.bytes_host(), s.block_off.to_numpy(), s.out_off.to_numpy(), v.size, False,
                                  nthreads=16)
    assert np.array_equal(gst, wst) and np.all(wst == 0)
    assert np.array_equal(gout, wout) and np.array_equal(gout, v)
    s.free()


@pytest.mark.parametrize("method", ["slot", "sized"])
@pytest.mark.parametrize("which", ["timestamps64", "uniform64"])
def test_u64_encode_bytes_and_every_block(ctx, orc, which, method):
    from utlpfor_amd import batch, workloads
    v = workloads.timestamps64(32 << 20) if which == "timestamps64" else workloads.uniform64(8 << 20)
    say(f"{which}: {v.size} values; oracle pack; GPU encode ({method})")
    want, bl, bc = orc.pack_stream(0, v, nthreads=16)
    s = batch.encode(ctx, v, flag=0, method=method)
    assert np.array_equal(s.meta.to_numpy(), _meta_of(bl, bc))
    assert np.array_equal(s.bytes_host(), want)
    say("decode of every block vs the oracle")
    out, st = batch.decode(ctx, s)
    ctx.sync()
    boff, ooff = s.block_off.to_numpy(), s.out_off.to_numpy()
    wout, wst = orc.unpack_blocks(want, boff, ooff, v.size, True, nthreads=16)
    gst = st.to_numpy()
    assert np.array_equal(gst, wst)
    good = np.repeat(wst == 0, 128)[:v.size]
    assert np.array_equal(out.to_numpy()[good], wout[good])
    # blocks the reference's decoder accepts round-trip to the values
    assert np.array_equal(wout[good], v[good])
    out.free()
    st.free()
    s.free()
