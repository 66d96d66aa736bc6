"""The Compressor32 / Compressor64 contract (compressor.go:16-107), restated
from the reference's compressor_test.go:12-330 on the GPU path: round trips
per flag, Append (BlockLength skips the first block), Append and NoInPlace
leave the caller's values unchanged, appended blocks with exceptions decode
without stale data, Get at every position, scratch (context) reuse across
rounds, a preallocated Decompress target that receives the values, and the
uint64 forms. Also the bytes of every Compress against the oracle's."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c32():
    from utlpfor_amd import codec
    return codec.NewUint32()


@pytest.fixture(scope="module")
def c64():
    from utlpfor_amd import codec
    return codec.NewUint64()


def _u32(v):
    return np.array(v, dtype=np.uint32)


def _u64(v):
    return np.array(v, dtype=np.uint64)


def _round_trip(c, orc, flag, values, is64=False):
    original = values.copy()
    packed, err = c.Compress(flag, None, values)
    assert err is None and len(packed) > 0
    want = (orc.PackUint64(flag, original)[0] if is64 else orc.PackUint32(flag, original)[0])
    assert packed == want
    unpacked, consumed, err = c.Decompress(None, packed)
    assert err is None and consumed == len(packed)
    assert np.array_equal(unpacked, original)
    return packed


def test_round_trip_raw(c32, orc):
    _round_trip(c32, orc, 0, _u32([10, 20, 30, 40, 50]))


def test_round_trip_delta(c32, orc):
    from utlpfor_amd import codec
    _round_trip(c32, orc, codec.Delta, _u32([100 + 3 * i for i in range(128)]))


def test_round_trip_nopatch_nofor(c32, orc):
    from utlpfor_amd import codec
    _round_trip(c32, orc, codec.NoPatch | codec.NoFOR, _u32([5, 12, 3, 7, 15, 1, 0, 8]))


def test_append(c32):
    from utlpfor_amd import codec
    block1, block2 = _u32([1, 2, 3, 4]), _u32([100, 200, 300, 400])
    orig2 = block2.copy()
    dst, err = c32.Compress(codec.Append, None, block1)
    assert err is None
    n1 = len(dst)
    dst, err = c32.Compress(codec.Append, dst, block2)
    assert err is None and len(dst) > n1
    bl, err = codec.BlockLength(dst)
    assert err is None and bl == n1
    unpacked, _, err = c32.Decompress(None, dst[bl:])
    assert err is None and np.array_equal(unpacked, orig2)


def test_append_preserves_values(c32):
    from utlpfor_amd import codec
    values = _u32([1000 + 7 * i for i in range(128)])
    original = values.copy()
    dst = None
    for _ in range(3):
        dst, err = c32.Compress(codec.Append, dst, values)
        assert err is None
    assert np.array_equal(values, original)


def test_noinplace_preserves_values(c32):
    from utlpfor_amd import codec
    values = _u32([1000 + 7 * i for i in range(128)])
    original = values.copy()
    _, err = c32.Compress(codec.NoInPlace, None, values)
    assert err is None and np.array_equal(values, original)


def test_inplace_work_values(c32, orc):
    """Without Append or NoInPlace the values slice receives the work values
    (FOR subtraction, delta), as PackUint32 leaves them."""
    from utlpfor_amd import codec
    values = _u32([1000 + 7 * i for i in range(128)])
    _, work, _ = orc.PackUint32(codec.Delta, values.copy())
    _, err = c32.Compress(codec.Delta, None, values)
    assert err is None and np.array_equal(values, work)


def test_append_with_exceptions(c32):
    from utlpfor_amd import codec

    def make_block(outlier):
        v = np.array([i % 16 for i in range(128)], dtype=np.uint32)
        v[::4] = outlier
        return v

    block1, block2 = make_block(0x10000), make_block(0x20000)
    orig1, orig2 = block1.copy(), block2.copy()
    probe, err = c32.Compress(0, None, block1.copy())
    assert err is None
    dst, err = c32.Compress(codec.Append, b"", block1)
    assert err is None
    dst, err = c32.Compress(codec.Append, dst, block2)
    assert err is None and len(dst) == 2 * len(probe)
    d1, consumed1, err = c32.Decompress(None, dst)
    assert err is None and np.array_equal(d1, orig1)
    d2, _, err = c32.Decompress(None, dst[consumed1:])
    assert err is None and np.array_equal(d2, orig2)


def test_get(c32):
    values = [100, 200, 300, 400, 500]
    packed, err = c32.Compress(0, None, _u32(values))
    assert err is None
    for i, want in enumerate(values):
        assert c32.Get(i, packed) == (want, None)


def test_get_delta(c32):
    from utlpfor_amd import codec
    values = _u32([1000 + 7 * i for i in range(64)])
    original = values.copy()
    packed, err = c32.Compress(codec.Delta, None, values)
    assert err is None
    for i, want in enumerate(original):
        assert c32.Get(i, packed) == (int(want), None)


def test_scratch_reuse(c32):
    from utlpfor_amd import codec
    for rnd in range(5):
        values = _u32([rnd * 1000 + i for i in range(32)])
        original = values.copy()
        packed, err = c32.Compress(codec.Delta, None, values)
        assert err is None
        unpacked, _, err = c32.Decompress(None, packed)
        assert err is None and np.array_equal(unpacked, original), rnd


def test_decompress_preallocated_target(c32):
    """A target with room for a block (capacity 128) receives the values and
    the result is a view of it (growUnpackDst, unpack_block.go:177-182); a
    smaller one is replaced by a new array, as in the reference."""
    values = _u32([10, 20, 30, 40, 50])
    original = values.copy()
    packed, err = c32.Compress(0, None, values)
    assert err is None
    target = np.full(128, 0xDEADBEEF, dtype=np.uint32)
    unpacked, consumed, err = c32.Decompress(target, packed)
    assert err is None and consumed == len(packed) and np.array_equal(unpacked, original)
    assert np.shares_memory(unpacked, target) and np.array_equal(target[:5], original)
    small = np.zeros(64, dtype=np.uint32)
    unpacked, _, err = c32.Decompress(small, packed)
    assert err is None and np.array_equal(unpacked, original) and not np.shares_memory(unpacked, small)


def test_full_block(c32, orc):
    from utlpfor_amd import codec
    _round_trip(c32, orc, codec.Delta, _u32([100 * i for i in range(128)]))


def test64_round_trip_raw(c64, orc):
    _round_trip(c64, orc, 0, _u64([1_000_000_000_000, 1_000_000_000_001, 1_000_000_000_002]), True)


def test64_round_trip_delta(c64, orc):
    from utlpfor_amd import codec
    _round_trip(c64, orc, codec.Delta, _u64([1_000_000_000_000 + 5 * i for i in range(128)]), True)


def test64_round_trip_small_values(c64, orc):
    _round_trip(c64, orc, 0, _u64([10, 20, 30, 40, 50]), True)


def test64_get(c64):
    values = [1_000_000_000_000, 2_000_000_000_000, 3_000_000_000_000]
    packed, err = c64.Compress(0, None, _u64(values))
    assert err is None
    for i, want in enumerate(values):
        assert c64.Get(i, packed) == (want, None)


def test64_scratch_reuse(c64):
    for rnd in range(5):
        values = _u64([rnd * 1_000_000_000_000 + i for i in range(32)])
        original = values.copy()
        packed, err = c64.Compress(0, None, values)
        assert err is None
        unpacked, _, err = c64.Decompress(None, packed)
        assert err is None and np.array_equal(unpacked, original) and np.array_equal(values, original), rnd


def test64_append(c64):
    from utlpfor_amd import codec
    block1, block2 = _u64([100, 200, 300]), _u64([1_000_000_000_000, 2_000_000_000_000])
    orig2 = block2.copy()
    dst, err = c64.Compress(codec.Append, None, block1)
    assert err is None
    n1 = len(dst)
    dst, err = c64.Compress(codec.Append, dst, block2)
    assert err is None and len(dst) > n1
    bl, err = codec.BlockLength(dst)
    assert err is None and bl == n1
    unpacked, _, err = c64.Decompress(None, dst[bl:])
    assert err is None and np.array_equal(unpacked, orig2)


def test64_decompress_target_by_count(c64):
    """uint64: a target reused when it holds the block's count
    (uint64.go:224-227)."""
    values = _u64([7, 8, 9])
    packed, err = c64.Compress(0, None, values)
    assert err is None
    target = np.zeros(3, dtype=np.uint64)
    unpacked, _, err = c64.Decompress(target, packed)
    assert err is None and np.shares_memory(unpacked, target) and np.array_equal(target, values)
