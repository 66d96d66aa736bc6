"""Stream container on the device: streams encoded on the GPU are written to
a container, read back, indexed by the device scan and decoded; the result
equals the source values, and the container's data section equals the
oracle's packed streams. A container built from a foreign-stream walk
decodes the same way."""
import os
import sys
import tempfile

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


@pytest.mark.parametrize("is64", [False, True])
def test_container_encode_save_load_decode(ctx, orc, is64):
    from utlpfor_amd import batch
    rng = np.random.default_rng(41 + is64)
    lens = np.array([1, 128, 129, 5000, 0, 777, 20000, 3], np.uint64)
    flags = np.array([0, 1, 2, 4, 0, 3, 1, 0], np.uint8)
    dt = np.uint64 if is64 else np.uint32
    top = 1 << 64 if is64 else 1 << 32
    parts = [rng.integers(0, top, int(n), dtype=np.uint64).astype(dt) if i % 2 else
             np.cumsum(rng.integers(0, 50, int(n))).astype(dt) for i, n in enumerate(lens)]
    v = np.concatenate(parts)
    say(f"container {'u64' if is64 else 'u32'}: {lens.size} streams, {v.size} values")
    s = batch.encode(ctx, v, stream_lengths=lens, stream_flags=flags)
    d = batch.directory(s, lens, flags)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "streams.utlc")
        n = batch.save_container(path, s, d)
        assert os.path.getsize(path) == n
        s2, d2 = batch.load_container(ctx, path)
    assert np.array_equal(d2, d) and s2.nblocks == s.nblocks and s2.nbytes == s.nbytes and s2.is64 == is64
    # data section: the oracle's Append stream of every stream, in order
    want = np.concatenate([orc.pack_stream(int(f), p)[0] if p.size else np.zeros(0, np.uint8)
                           for f, p in zip(flags, parts)])
    assert np.array_equal(s2.bytes_host(), want)
    out, st = batch.decode(ctx, s2)
    ctx.sync()
    assert np.all(st.to_numpy() == 0) and np.array_equal(out.to_numpy(), v)
    # per-stream access through the directory
    got = out.to_numpy()
    for e, p in zip(d2, parts):
        a = int(e["first_value"])
        assert np.array_equal(got[a:a + int(e["nvalues"])], p)
    s.free()
    s2.free()


def test_container_from_foreign_walk(ctx, orc):
    from utlpfor_amd import batch
    from utlpfor_amd.device import DeviceArray
    rng = np.random.default_rng(43)
    parts = [np.cumsum(rng.integers(0, 9, int(n))).astype(np.uint32) for n in (700, 40000, 1, 3000)]
    streams = [orc.pack_stream(1, p)[0] for p in parts]
    blob = np.concatenate(streams)
    off = np.concatenate([[0], np.cumsum([x.size for x in streams])]).astype(np.uint64)
    data = DeviceArray.from_numpy(np.concatenate([blob, np.zeros(16, np.uint8)]), ctx)
    s, counts, status = batch.index_foreign(ctx, data, off)
    assert np.all(status == 0)
    blk = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    d = np.zeros(len(parts), batch.STREAM_ENTRY)
    d["first_block"], d["nblocks"] = blk[:-1], counts
    d["nvalues"] = [p.size for p in parts]
    d["first_value"] = np.concatenate([[0], np.cumsum(d["nvalues"])[:-1]])
    d["byte_off"], d["flag"] = off[:-1], 1
    buf = batch.to_container(s, d)
    s2, d2 = batch.from_container(ctx, buf)
    out, st = batch.decode(ctx, s2)
    ctx.sync()
    assert np.all(st.to_numpy() == 0) and np.array_equal(out.to_numpy(), np.concatenate(parts))
    s2.free()
