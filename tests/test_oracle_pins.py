"""Pins the CPU restatement (the oracle) before anything is compared to it.

Each check is a fixture, known answer or invariant taken from the
reference's own tests, or a property the mathematics fixes, chosen so that
a plausible restatement mistake (dropped term, wrong index, swapped
operand) fails at least one of them:

* golden blocks            format_conformance_test.go:693-754 (testdata/*.bin)
* Example sizes            example_test.go:10-340
* MaxBlockLength           max_block_length_test.go:12-44, example_test.go:302-330
* BlockLength table/errs   block_length_test.go:37-59, :127-176
* stream walk              example_test.go:171-187 (BlockLength skip), packed
                           block lengths, truncation and header stops
* overflow position        delta_test.go:229-244, simd_delta_overflow_test.go:177-215
* FOR bases and widths     for_test.go:110-175, :457-478
* FOR64 layout             uint64_test.go:1143-1214
* in-place mutation        delta_test.go:187-200
* corrupt corpus/sweeps    corrupt_test.go:170-378, testdata/fuzz/*
* layout vs spec           tests/wire_spec.py (independent reader of utlpfor.ksy)
* cost model               brute-force enumeration of all 9 candidate widths

Parity unpinned (no fixture or closed form reaches it): exact bytes of
StreamVByte 4-byte codes (pinned only by the Stream VByte definition and
the decoder's LUT, svb.go:9-24), and the uint64 two-block quirk where the
upper sub-block gets a 4-byte FOR base (derived from uint64.go:174 and
unpack_block.go:99-109; untested in the reference).
"""
import os
import struct

import numpy as np
import pytest

from tests import wire_spec

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ALL_FLAGS = [f for f in range(16)]  # Delta|NoFOR|NoPatch|Special combos
DOC_ERRORS = {0, 1, 2, 3, 4, 5, 6}


def golden(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def golden_inputs():
    exc = list(range(128))
    for p in (5, 15, 25, 35, 45, 55, 65, 75, 85, 95):
        exc[p] = 0x10000000 + p
    return [
        ("plain_128_bw8.bin", [128 + i for i in range(128)], 0),
        ("plain_128_bw0.bin", [0] * 128, 0),
        ("delta_128.bin", [100 * i for i in range(128)], 1),
        ("exceptions_128.bin", exc, 0),
    ]


@pytest.mark.parametrize("name,values,flag", golden_inputs())
def test_golden_vectors(orc, name, values, flag):
    packed, _, err = orc.PackUint32(flag, values)
    assert err == 0
    assert packed == golden(name)
    out, consumed, err, _ = orc.UnpackUint32(packed)
    assert err == 0 and list(out) == values
    bl, err = orc.BlockLength(packed)
    assert err == 0 and bl == consumed == len(packed)
    for pos, want in enumerate(values):
        assert orc.GetUint32(pos, packed) == (want, 0)
    assert wire_spec.parse_u32_block(packed)[0] == values


def test_example_sizes(orc):
    P = lambda f, v: orc.PackUint32(f, v)[0]
    assert len(P(0, [10, 20, 30, 40, 50])) == 18
    v = [100000 + 4 * i for i in range(128)]
    assert (len(P(0, v)), len(P(orc.Delta, v))) == (200, 136)
    v = [4_000_000 + i % 16 for i in range(128)]
    v[10], v[90] = 4_500_000, 4_600_000
    assert [len(P(f, v)) for f in (0, orc.NoFOR, orc.NoPatch, orc.NoFOR | orc.NoPatch)] == [81, 388, 328, 388]
    assert (len(P(orc.Append, [1, 2, 3, 4])), len(P(orc.Append, [100, 200, 300, 400]))) == (15, 17)
    assert 3 * len(P(orc.Append, [500 + 3 * i for i in range(128)])) == 588
    assert len(P(orc.Append, [1, 2, 3])) == 13
    assert orc.Header(P(orc.Delta | orc.NoPatch, [10 * i for i in range(128)])) == (
        128, 8, 0, True, False, False, False, 0)
    p64, err = orc.PackUint64(0, [10**12, 10**12 + 1, 10**12 + 2])
    assert err == 0 and len(p64) == 19
    out, consumed, err, _ = orc.UnpackUint64(p64)
    assert list(out) == [10**12, 10**12 + 1, 10**12 + 2] and consumed == 19
    assert orc.GetUint64(1, p64) == (10**12 + 1, 0)
    p64, _ = orc.PackUint64(0, [10**12, 2 * 10**12, 3 * 10**12])
    assert orc.GetUint64(1, p64) == (2 * 10**12, 0)
    v = [1000 + 4 * i for i in range(128)]
    p = P(orc.Delta, v)
    assert len(p) == 170 and list(orc.UnpackUint32(p)[0][:5]) == [1000, 1004, 1008, 1012, 1016]
    p = P(0, [100, 200, 300, 400, 500])
    assert orc.GetUint32(2, p) == (300, 0)
    # example_test.go:171-187: skip over the first block by BlockLength
    b1, b2 = P(orc.Append, [1, 2, 3]), P(orc.Append, [100, 200, 300])
    buf = b1 + b2
    bl, _ = orc.BlockLength(buf)
    assert bl == 13 and list(orc.UnpackUint32(buf[bl:])[0]) == [100, 200, 300]


def test_max_block_length(orc):
    assert [orc.MaxBlockLength32(f) for f in (0, orc.NoPatch, orc.NoPatch | orc.NoFOR, orc.NoFOR)] == [
        1018, 520, 516, 1014]
    assert [orc.MaxBlockLength64(f) for f in (0, orc.NoPatch, orc.NoPatch | orc.NoFOR, orc.NoFOR)] == [
        2038, 1042, 1034, 2030]
    # Delta, Special and Append do not change the bound
    assert orc.MaxBlockLength32(orc.Delta | orc.Special | orc.Append) == 1018


def _hdr(count, bw, exc, flags):
    return struct.pack("<I", (count & 0xFF) | ((bw // 4) << 8) | flags | (exc << 24))


T32, T64 = 2 << 13, 3 << 13


def test_block_length_table(orc):
    for count, bw, want in [(128, 0, 4), (128, 4, 68), (128, 8, 132), (128, 16, 260), (128, 32, 516),
                            (50, 8, 132)]:
        buf = _hdr(count, bw, 0, T32) + b"\0" * 1024
        assert orc.BlockLength(buf) == (want, 0)
    for flags in (T32 | 1 << 18, T32 | 1 << 19, T32 | 1 << 20, T64 | 1 << 18, T64 | 1 << 20):
        assert orc.BlockLength(_hdr(128, 8, 0, flags) + b"\0" * 1024)[1] == orc.ERR_INVALID_FLAGS
    bad_bw = struct.pack("<I", 128 | (9 << 8) | T32) + b"\0" * 1024
    assert orc.BlockLength(bad_bw)[1] == orc.ERR_INVALID_BUFFER
    assert orc.BlockLength(b"\x01\x02\x03")[1] == orc.ERR_INVALID_BUFFER
    assert orc.BlockLength(_hdr(129, 8, 0, T32))[1] == orc.ERR_INVALID_BLOCK_LENGTH
    # exceptions: header + svbLen + payload + index + svb
    buf = _hdr(128, 8, 3, T32) + struct.pack("<H", 7) + b"\0" * 600
    assert orc.BlockLength(buf) == (4 + 2 + 128 + 3 + 7, 0)
    buf = _hdr(128, 8, 40, T32) + struct.pack("<H", 50) + b"\0" * 600
    assert orc.BlockLength(buf) == (4 + 2 + 128 + 16 + 50, 0)


def _delta_block(deltas, count=128):
    """A delta block whose stored per-lane deltas are `deltas` (no zigzag)."""
    raw, _, err = oracle_mod().PackUint32(2 | 4, deltas[:count])  # NoFOR|NoPatch keeps the raw words
    assert err == 0
    h = struct.unpack_from("<I", raw)[0] | (1 << 22)
    return struct.pack("<I", h) + raw[4:]


def oracle_mod():
    from oracle import oracle
    return oracle


def test_overflow_position(orc):
    d = [0] * 128
    d[0], d[16] = 0xFFFFFFF0, 0x20
    out, _, err, pos = orc.UnpackUint32(_delta_block(d))
    assert err == orc.ERR_OVERFLOW and pos == 16
    # every (row, lane): closed form row*16 + lane
    for row in range(1, 8):
        for lane in range(16):
            d = [0] * 128
            d[lane], d[row * 16 + lane] = 0xFFFFFFF0, 0x20
            _, _, err, pos = orc.UnpackUint32(_delta_block(d))
            assert (err, pos) == (orc.ERR_OVERFLOW, row * 16 + lane)
    # v-major: the lowest row wins over a lower lane in a later row
    d = [0] * 128
    d[0], d[2 * 16 + 0] = 0xFFFFFFF0, 0x20
    d[9], d[1 * 16 + 9] = 0xFFFFFFF0, 0x20
    assert orc.UnpackUint32(_delta_block(d))[2:] == (orc.ERR_OVERFLOW, 25)
    # partial block: a wrap at or beyond count is not reported
    d = [0] * 128
    d[3], d[16 + 3] = 0xFFFFFFF0, 0x20
    assert orc.UnpackUint32(_delta_block(d, 20))[2:] == (orc.ERR_OVERFLOW, 19)
    out, _, err, _ = orc.UnpackUint32(_delta_block(d, 19))
    assert err == 0 and len(out) == 19


def test_for_bases(orc):
    def hdr_base(p):
        h = struct.unpack_from("<I", p)[0]
        fw = (h >> 15) & 3
        off = 4 + (2 if h >> 24 else 0)
        return fw, int.from_bytes(p[off:off + {0: 0, 1: 1, 2: 2, 3: 4}[fw]], "little")
    assert hdr_base(orc.PackUint32(0, [1000000 + i % 100 for i in range(128)])[0]) == (3, 1000000)
    fw, base = hdr_base(orc.PackUint32(0, [200 + i % 10 for i in range(128)])[0])
    assert fw == 0 or (fw, base) == (1, 200)
    fw, base = hdr_base(orc.PackUint32(0, [50000 + i % 10 for i in range(128)])[0])
    assert fw == 0 or (fw, base) == (2, 50000)
    assert hdr_base(orc.PackUint32(0, list(range(128)))[0])[0] == 0
    assert hdr_base(orc.PackUint32(0, [i * 100000 for i in range(128)])[0])[0] == 0
    assert hdr_base(orc.PackUint32(0, [999999] * 128)[0]) == (3, 999999)
    # selectFORWidth table through single-value blocks (min == max == v: FOR
    # chosen whenever the raw step is > 0)
    for v, want in [(1, 1), (255, 1), (256, 2), (65535, 2), (65536, 3), (0xFFFFFFFF, 3)]:
        assert hdr_base(orc.PackUint32(0, [v] * 128)[0]) == (want, v)


def test_for64_layout(orc):
    base = 0x100000000
    p, _ = orc.PackUint64(0, [base + i for i in range(128)])
    assert (struct.unpack_from("<I", p)[0] >> 15) & 3 == 3
    p, _ = orc.PackUint64(orc.NoPatch, [base + i for i in range(128)])
    assert struct.unpack_from("<Q", p, 4)[0] == base
    v = [base + i for i in range(128)]
    v[0] = base + 0xFFFFF
    p, _ = orc.PackUint64(0, v)
    h = struct.unpack_from("<I", p)[0]
    assert h >> 24 > 0 and struct.unpack_from("<Q", p, 6)[0] == base + 1
    v = [base + 7 * i for i in range(128)]
    p, _ = orc.PackUint64(0, v)
    assert list(orc.UnpackUint64(p)[0]) == v


def test_two_block_quirk_upper_for32(orc):
    """uint64.go:174 lets the upper sub-block use a 4-byte FOR base; the
    decoder then reads it as FOR64 (unpack_block.go:99-109) and rejects."""
    v = [0x1000000000000000 + ((i % 4) << 32) + i * 12345 for i in range(128)]
    p, err = orc.PackUint64(0, v)
    assert err == 0
    h1 = struct.unpack_from("<I", p)[0]
    assert h1 & (1 << 19)
    assert orc.UnpackUint64(p)[2] == orc.ERR_INVALID_BUFFER
    assert orc.GetUint64(0, p)[1] == orc.ERR_INVALID_BUFFER


def test_in_place_mutation(orc):
    values = [10 * i for i in range(128)]
    _, work, err = orc.PackUint32(orc.Delta, values)
    assert work[0] == 0 and work[16] == 160 and list(work) != values
    _, work, _ = orc.PackUint32(orc.Delta | orc.NoInPlace, values)
    assert list(work) == values
    _, work, _ = orc.PackUint32(orc.Delta | orc.Append, values)
    assert list(work) == values


def _zz(x):
    x &= 0xFFFFFFFF
    s = x - (1 << 32) if x >> 31 else x
    return ((s << 1) ^ (s >> 31)) & 0xFFFFFFFF


def _cost_bruteforce(work):
    """Enumerate every step width directly (no suffix sums) and return the
    width the cost model picks: strictly smaller cost wins, ties keep the
    wider width (select_bitwidth.go:41-121)."""
    orall = 0
    for w in work:
        orall |= w
    max_si = (orall.bit_length() + 3) >> 2
    best = (4 + 64 * max_si, max_si * 4, 0)
    bpv = [1, 1, 1, 2, 2, 3, 3, 4, 4]
    for si in range(max_si - 1, -1, -1):
        exc = sum(1 for w in work if w.bit_length() > 4 * si)
        cost = 4 + 64 * si + 2 + min(exc, 16) + (exc + 3) // 4 + exc * bpv[max_si - si]
        if cost < best[0]:
            best = (cost, 4 * si, exc)
    return best[1], best[2]


def _random_blocks(seed, n_blocks):
    rng = np.random.default_rng(seed)
    for _ in range(n_blocks):
        count = int(rng.choice([1, 2, 3, 5, 15, 16, 17, 31, 64, 100, 127, 128]))
        kind = rng.integers(0, 6)
        if kind == 0:
            v = rng.integers(0, 1 << int(rng.integers(0, 33)), count, dtype=np.uint64)
        elif kind == 1:
            v = np.cumsum(rng.integers(0, 1000, count)) + int(rng.integers(0, 1 << 31))
        elif kind == 2:
            v = rng.integers(0, 1 << 12, count, dtype=np.uint64)
            m = rng.random(count) < 0.1
            v[m] = rng.integers(1 << 20, 1 << 32, int(m.sum()), dtype=np.uint64)
        elif kind == 3:
            v = np.cumsum(rng.integers(-1024, 1025, count)) + 5000
        elif kind == 4:
            v = rng.integers(0, 1 << 32, count, dtype=np.uint64)
        else:
            v = np.full(count, int(rng.integers(0, 1 << 32)))
        yield [int(x) & 0xFFFFFFFF for x in v]


def test_roundtrip_and_spec_all_flags(orc):
    for i, values in enumerate(_random_blocks(1, 240)):
        flag = ALL_FLAGS[i % 16]
        packed, work, err = orc.PackUint32(flag, values)
        assert err == 0
        assert len(packed) <= orc.MaxBlockLength32(flag)
        out, consumed, err, _ = orc.UnpackUint32(packed)
        assert err == 0 and list(out) == values
        assert orc.BlockLength(packed) == (len(packed), 0) and consumed == len(packed)
        assert wire_spec.parse_u32_block(packed)[0] == values
        for pos in range(len(values)):
            assert orc.GetUint32(pos, packed) == (values[pos], 0)
        nip, work2, _ = orc.PackUint32(flag | orc.NoInPlace, values)
        assert nip == packed and list(work2) == values
        h = struct.unpack_from("<I", packed)[0]
        assert bool(h & (1 << 17)) == bool(flag & orc.Special)
        if not flag & orc.NoPatch:
            bw, exc = (h >> 8 & 31) * 4, h >> 24
            assert (bw, exc) == _cost_bruteforce([int(w) for w in work])


def test_zigzag_closed_form(orc):
    values = [(1000 + 7 * i) if (i // 16) % 2 == 0 else (500 + i) for i in range(128)]
    packed, work, err = orc.PackUint32(orc.Delta | orc.NoFOR, values)
    h = struct.unpack_from("<I", packed)[0]
    assert h & (1 << 23)
    for i in range(128):
        d = values[i] if i < 16 else (values[i] - values[i - 16])
        assert int(work[i]) == _zz(d)
    assert list(orc.UnpackUint32(packed)[0]) == values


def test_tiny_bruteforce(orc):
    import itertools
    pool = [0, 1, 15, 16, 255, 256, 65535, 1 << 24, 0xFFFFFFFF]
    for count in (1, 2, 3):
        for values in itertools.product(pool, repeat=count):
            for flag in range(16):
                packed, _, err = orc.PackUint32(flag, list(values))
                assert err == 0
                out, _, err, _ = orc.UnpackUint32(packed)
                assert err == 0 and tuple(int(x) for x in out) == values
                assert wire_spec.parse_u32_block(packed)[0] == list(values)


def test_u64_roundtrip(orc):
    rng = np.random.default_rng(5)
    cases = [
        [int(x) for x in rng.integers(0, 1 << 63, 128, dtype=np.uint64) * 2 + 1],
        [int(1.7e18) + int(x) for x in np.cumsum(rng.geometric(1 / 1000, 100))],
        [int(x) for x in rng.integers(0, 1 << 20, 77)],
        [0x1234567800000000 | (i * 7) for i in range(128)],
    ]
    for values in cases:
        for flag in range(16):
            p, err = orc.PackUint64(flag, values)
            assert err == 0 and len(p) <= orc.MaxBlockLength64(flag)
            out, consumed, err, _ = orc.UnpackUint64(p)
            if err == orc.ERR_INVALID_BUFFER:
                continue  # the two-block quirk (test_two_block_quirk_upper_for32)
            assert err == 0 and list(out) == values and consumed == len(p)
            assert wire_spec.parse_u64_block(p)[0] == values
            assert orc.BlockLength(p) == (len(p), 0)
            for pos in (0, len(values) // 2, len(values) - 1):
                assert orc.GetUint64(pos, p) == (values[pos], 0)


def test_error_precedence_examples(orc):
    assert orc.UnpackUint32(b"")[2] == orc.ERR_INVALID_BUFFER
    assert orc.UnpackUint32(_hdr(5, 8, 0, 3 << 13))[2] == orc.ERR_UNSUPPORTED_TYPE
    out, consumed, err, _ = orc.UnpackUint32(_hdr(0, 36, 0, T32))
    assert err == 0 and consumed == 4 and len(out) == 0
    assert orc.UnpackUint32(_hdr(200, 0, 0, T32))[2] == orc.ERR_INVALID_BLOCK_LENGTH
    assert orc.UnpackUint32(struct.pack("<I", 5 | (9 << 8) | T32))[2] == orc.ERR_INVALID_BUFFER
    assert orc.GetUint32(5, _hdr(5, 8, 0, T32) + b"\0" * 200)[1] == orc.ERR_POSITION_OUT_OF_RANGE
    assert orc.GetUint32(-1, _hdr(5, 8, 0, T32) + b"\0" * 200)[1] == orc.ERR_POSITION_OUT_OF_RANGE
    # excCount > count
    assert orc.UnpackUint32(_hdr(2, 8, 3, T32) + struct.pack("<H", 4) + b"\0" * 200)[2] == 1
    # Get accepts count > 128 on non-delta blocks (no count check, get.go:102)
    v, err = orc.GetUint32(200, _hdr(250, 8, 0, T32) + b"\x11" * 200)
    assert err == 0 and v == 0
    # ... and the reference panics on delta blocks at rows 8 (pos 128..143)
    assert orc.GetUint32(130, _hdr(250, 8, 0, T32 | 1 << 22) + b"\0" * 200)[1] == orc.REF_PANIC
    assert orc.GetUint32(150, _hdr(250, 8, 0, T32 | 1 << 22) + b"\0" * 200)[1] == orc.ERR_INVALID_BLOCK_LENGTH
    # Header rejects uint64 blocks, BlockLength does not check the type
    assert orc.Header(_hdr(5, 8, 0, T64))[-1] == orc.ERR_UNSUPPORTED_TYPE
    assert orc.BlockLength(_hdr(5, 8, 0, 0))[1] == 0


def _corrupt_targets(orc):
    fill = lambda f: [f(i) for i in range(128)]
    rng = np.random.default_rng(7)
    few = fill(lambda i: i % 50)
    few[7], few[50] = 0xFFFF0000, 0x00FF0000
    many = fill(lambda i: i % 50)
    for j in range(48):
        many[(j * 11) % 128] = 0xFF000000 | int(rng.integers(0, 0xFFFFFF))
    dexc = fill(lambda i: 1000 + 3 * i)
    dexc[64] = 0xFFFF0000
    defs = [(0, fill(lambda i: i % 200)), (0, few), (0, many), (1, fill(lambda i: 1000 + 3 * i)),
            (1, fill(lambda i: 1000000 + 3 * i)),
            (1, fill(lambda i: 1000 + 3 * i if i % 2 == 0 else 1000 + 3 * (i - 1) - 97)), (1, dexc)]
    t32 = [orc.PackUint32(f, v)[0] for f, v in defs]
    t64 = [orc.PackUint64(0, fill(f))[0] for f in (lambda i: i % 200, lambda i: 0x100000000 + 3 * i,
                                                    lambda i: 0x1234567800000000 | (i * 7))]
    return t32, t64


def test_corrupt_sweeps_documented_errors(orc):
    t32, t64 = _corrupt_targets(orc)
    positions = [0, 1, 7, 15, 16, 31, 64, 100, 127]
    for packed in t32:
        for cut in range(1, len(packed)):
            tr = packed[:cut]
            for pos in positions:
                got, err = orc.GetUint32(pos, tr)
                assert err in DOC_ERRORS
                if err == 0:
                    assert (got, 0) == orc.GetUint32(pos, packed)
        h = struct.unpack_from("<I", packed)[0]
        if h >> 24:
            bad = bytearray(packed)
            bad[4:6] = b"\xff\xff"
            for pos in positions:
                assert orc.GetUint32(pos, bytes(bad))[1] == orc.ERR_INVALID_BUFFER
    for packed in t32 + t64:
        for i in range(len(packed)):
            for probe in (0x00, 0xFF):
                bad = bytearray(packed)
                bad[i] = probe
                bad = bytes(bad)
                for pos in positions:
                    assert orc.GetUint32(pos, bad)[1] in DOC_ERRORS
                    assert orc.GetUint64(pos, bad)[1] in DOC_ERRORS
                assert orc.UnpackUint32(bad)[2] in DOC_ERRORS
                assert orc.UnpackUint64(bad)[2] in DOC_ERRORS
                assert orc.BlockLength(bad)[1] in DOC_ERRORS
                assert orc.Header(bad)[-1] in DOC_ERRORS


def test_fuzz_corpus(orc):
    corpus = [b"\x00a800", b"0A10", b"0A00\x0e\x00" + b"0" * 64 + b"1" + b"0" * 16 + b"2" + b"0" * 81]
    for data in corpus:
        for pos in (0, 7, 64, 127):
            assert orc.GetUint32(pos, data)[1] in DOC_ERRORS
            assert orc.GetUint64(pos, data)[1] in DOC_ERRORS
        assert orc.UnpackUint32(data)[2] in DOC_ERRORS
        assert orc.UnpackUint64(data)[2] in DOC_ERRORS
        assert orc.BlockLength(data)[1] in DOC_ERRORS
        assert orc.Header(data)[-1] in DOC_ERRORS


def test_stream_helpers_match_per_block(orc):
    rng = np.random.default_rng(3)
    values = rng.integers(0, 1 << 12, 1000, dtype=np.uint32)
    stream, bl, bc = orc.pack_stream(0, values, nthreads=3)
    assert bc.sum() == 1000 and bl.sum() == stream.size
    off = np.concatenate([[0], np.cumsum(bl)]).astype(np.uint64)
    out_off = np.concatenate([[0], np.cumsum(bc)]).astype(np.uint64)
    out, st = orc.unpack_blocks(stream, off, out_off, 1000, False, nthreads=4)
    assert (st == 0).all() and (out == values).all()
    for k in range(bl.size):
        blk, _, _ = orc.PackUint32(orc.Append, values[128 * k:128 * k + 128])
        assert blk == stream[int(off[k]):int(off[k + 1])].tobytes()


def test_walk_stream(orc):
    """The serial stream walk: the example's two-block skip
    (example_test.go:171-187), block starts equal to the scanned packed
    lengths, a per-block BlockLength loop, and its three stops."""
    P = lambda f, v: orc.PackUint32(f, v)[0]
    buf = np.frombuffer(P(orc.Append, [1, 2, 3]) + P(orc.Append, [100, 200, 300]), np.uint8)
    st, ln, ct, status = orc.walk_stream(buf)
    assert list(st) == [0, 13] and list(ln) == [13, buf.size - 13] and list(ct) == [3, 3] and status == 0
    rng = np.random.default_rng(3)
    for flag in (0, orc.Delta, orc.NoPatch, orc.Delta | orc.NoFOR):
        v = rng.integers(0, 1 << int(rng.integers(1, 33)), 5000, dtype=np.uint64).astype(np.uint32)
        stream, bl, bc = orc.pack_stream(flag, v)
        st, ln, ct, status = orc.walk_stream(stream)
        assert status == 0 and np.array_equal(ln, bl) and np.array_equal(ct, bc)
        assert np.array_equal(st, np.concatenate([[0], np.cumsum(bl.astype(np.uint64))[:-1]]))
        # the same walk, one BlockLength call per block
        b = stream.tobytes()
        pos, k = 0, 0
        while pos < len(b):
            L, err = orc.BlockLength(b[pos:])
            assert err == 0 and st[k] == pos and ln[k] == L
            pos, k = pos + L, k + 1
        assert k == st.size
        # stops: a truncated last block, a reserved header bit, a short tail
        st2, _, _, s2 = orc.walk_stream(stream[:-1])
        assert s2 == orc.ERR_INVALID_BUFFER and st2.size == st.size - 1
        bad = stream.copy()
        bad[int(st[7]) + 2] |= 0x04
        st3, _, _, s3 = orc.walk_stream(bad)
        assert s3 == orc.ERR_INVALID_FLAGS and st3.size == 7
        st4, _, _, s4 = orc.walk_stream(np.concatenate([stream, np.zeros(3, np.uint8)]))
        assert s4 == orc.ERR_INVALID_BUFFER and st4.size == st.size
    assert orc.walk_stream(np.zeros(0, np.uint8))[3] == 0
