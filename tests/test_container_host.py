"""The stream container (utl_container_*): host-only, no device work.

Streams are packed by the oracle (the reference's Append pattern,
pack.go:53-76), written through the C-ABI, read back by the C-ABI parser
and by an independent reader of the documented layout, and every kind of
structural damage is rejected.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from utlpfor_amd import _lib, batch


def _streams(orc, rng):
    parts = []
    for i, n in enumerate([0, 1, 127, 128, 129, 1000, 4096, 300]):
        is64 = i % 3 == 2
        if is64:
            v = rng.integers(0, 1 << 64, n, dtype=np.uint64)
        else:
            v = rng.integers(0, 1 << int(rng.integers(1, 33)), n, dtype=np.uint64).astype(np.uint32)
        flag = [0, 1, 2, 4][i % 4]
        if n:
            data, bl, bc = orc.pack_stream(flag, v)
        else:
            data, bl, bc = np.zeros(0, np.uint8), np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        parts.append((data, bl, bc, n, flag, is64))
    return parts


def _assemble(parts):
    d = np.zeros(len(parts), batch.STREAM_ENTRY)
    b = v = o = 0
    for i, (data, bl, bc, n, flag, is64) in enumerate(parts):
        d[i] = (b, bl.size, v, n, o, flag, int(is64))
        b += bl.size
        v += n
        o += data.size
    meta = np.concatenate([(p[1] & 0xFFFFFF) | (p[2] << 24) for p in parts]).astype(np.uint32)
    data = np.concatenate([p[0] for p in parts]).astype(np.uint8)
    return d, meta, data


def _write(d, meta, data):
    lib = _lib.load()
    n = lib.utl_container_bytes(d.size, meta.size, data.size)
    out = np.zeros(n, np.uint8)
    out_len = C.c_size_t()
    code = lib.utl_container_write(out.ctypes.data, n, d.ctypes.data, d.size, meta.ctypes.data, meta.size,
                                   data.ctypes.data, data.size, C.byref(out_len))
    return code, out, out_len.value


def _parse(buf):
    lib = _lib.load()
    buf = np.ascontiguousarray(buf)
    hdr = np.zeros(1, batch.CONTAINER_HEADER)
    p = [C.c_void_p() for _ in range(3)]
    return lib.utl_container_parse(buf.ctypes.data, buf.size, hdr.ctypes.data, *[C.byref(x) for x in p])


def test_container_roundtrip_and_layout(orc):
    rng = np.random.default_rng(4)
    parts = _streams(orc, rng)
    d, meta, data = _assemble(parts)
    code, buf, n = _write(d, meta, data)
    assert code == 0 and n == buf.size
    h, dir_, m, x = batch.parse_container(buf)
    assert np.array_equal(dir_, d) and np.array_equal(m, meta) and np.array_equal(x, data)
    # an independent reader of the documented layout
    magic, ver, hb, ns, nb, nx, nv, do, mo, xo, tot = struct.unpack_from("<8sII8Q", buf.tobytes(), 0)
    assert magic == b"UTLPFOR1" and ver == 1 and hb == 80
    assert (ns, nb, nx, nv) == (len(parts), meta.size, data.size, sum(p[3] for p in parts))
    assert do == 80 and mo == 80 + 48 * ns and xo == mo + (4 * nb + 7) // 8 * 8 and tot == xo + nx == buf.size
    for i, p in enumerate(parts):
        fb, nbk, fv, nvs, bo, flag, is64 = struct.unpack_from("<5Q2I", buf.tobytes(), do + 48 * i)
        assert (nbk, nvs, flag, is64) == (p[1].size, p[3], p[4], int(p[5]))
        # the stream's bytes are the packed blocks, byte for byte
        assert bytes(buf[xo + bo:xo + bo + p[0].size]) == p[0].tobytes()
        # and its meta words give the reference's block lengths in order
        mw = np.frombuffer(buf.tobytes(), "<u4", nbk, mo + 4 * fb)
        assert np.array_equal(mw & 0xFFFFFF, p[1]) and np.array_equal(mw >> 24, p[2])
    # an empty container
    code, buf0, n0 = _write(np.zeros(0, batch.STREAM_ENTRY), np.zeros(0, np.uint32), np.zeros(0, np.uint8))
    assert code == 0 and n0 == 80 and _parse(buf0) == 0


def test_container_rejects_damage(orc):
    rng = np.random.default_rng(5)
    d, meta, data = _assemble(_streams(orc, rng))
    code, buf, _ = _write(d, meta, data)
    assert code == 0
    INVALID = 1
    bad = buf.copy(); bad[0] ^= 1
    assert _parse(bad) == INVALID  # magic
    bad = buf.copy(); bad[8] = 2
    assert _parse(bad) == INVALID  # version
    assert _parse(buf[:-1]) == INVALID  # truncated
    assert _parse(buf[:79]) == INVALID  # short header
    h = batch.CONTAINER_HEADER
    for field in ("nstreams", "nblocks", "nbytes", "nvalues", "dir_off", "meta_off", "data_off", "total_bytes"):
        bad = buf.copy()
        off = h.fields[field][1]
        bad[off:off + 8] = np.frombuffer(struct.pack("<Q", (1 << 62) + 3), np.uint8)
        assert _parse(bad) == INVALID, field
    mo = 80 + 48 * d.size
    bad = buf.copy(); bad[mo] ^= 1  # a block length no longer sums to its stream's bytes
    assert _parse(bad) == INVALID
    bad = buf.copy(); bad[mo + 3] = 0  # value count 0
    assert _parse(bad) == INVALID
    bad = buf.copy(); bad[mo + 3] = 200  # value count > 128
    assert _parse(bad) == INVALID
    bad = buf.copy(); bad[80 + 48 + 8] ^= 1  # directory no longer tiles the blocks
    assert _parse(bad) == INVALID
    bad = buf.copy(); bad[80 + 48 * 2 + 44] = 5  # is_u64 not 0/1
    assert _parse(bad) == INVALID
    # the writer refuses an inconsistent directory and reports the size needed
    d2 = d.copy(); d2["nvalues"][1] += 1
    assert _write(d2, meta, data)[0] == 100
    lib = _lib.load()
    need = C.c_size_t()
    assert lib.utl_container_write(None, 0, d.ctypes.data, d.size, meta.ctypes.data, meta.size, data.ctypes.data,
                                   data.size, C.byref(need)) == 100 and need.value == buf.size
    assert lib.utl_container_bytes(1 << 62, 1 << 62, 1) == 0  # overflow
