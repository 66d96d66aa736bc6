"""Go-shaped per-call API (the reference's public surface, README.md:25-72).

Each function makes exactly one C-ABI call, which runs one GPU kernel for
Pack/Unpack/Get (Header, BlockLength and MaxBlockLength are host reads of a
few bytes, as in the reference). Results come back as Go does: a value or
slice, and an error that is one of the sentinels below or an ErrOverflow.

Reference: utlpfor.go:21-115 (flags, errors), pack.go:76, unpack.go:5,
get.go:74, block_length.go:8, header.go:135, scalar.go:336, uint64.go:279,
:338, :364, :451, compressor.go:16-107.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import Context

# ---- flags (utlpfor.go:21-68) ----
Delta, NoFOR, NoPatch, Special, Append, NoInPlace = 1, 2, 4, 8, 16, 32
ScratchLen, ScratchLenNoInPlace, ScratchLen64 = 128, 256, 384
IntTypeUint8, IntTypeUint16, IntTypeUint32, IntTypeUint64 = 0, 1, 2, 3


class UtlError(Exception):
    def __init__(self, msg: str, code: int):
        super().__init__(msg)
        self.code = code


ErrInvalidBuffer = UtlError("UTLpfor: invalid buffer", 1)
ErrPositionOutOfRange = UtlError("UTLpfor: position out of range", 2)
ErrInvalidBlockLength = UtlError("UTLpfor: invalid block length", 3)
ErrInvalidFlags = UtlError("UTLpfor: invalid header flags", 4)
ErrUnsupportedType = UtlError("UTLpfor: unsupported integer type", 5)
ErrNotImplemented = UtlError("UTLpfor: not implemented", 7)


class ErrOverflow(UtlError):
    """utlpfor.go:87-115: Position = row*16 + lane of the first wrap."""

    def __init__(self, position: int):
        super().__init__(f"UTLpfor: delta decode overflow at index {position}", 6)
        self.Position = position

    def __eq__(self, other):
        return isinstance(other, ErrOverflow) and other.Position == self.Position

    def __hash__(self):
        return hash(("overflow", self.Position))


_BY_CODE = {e.code: e for e in (ErrInvalidBuffer, ErrPositionOutOfRange, ErrInvalidBlockLength, ErrInvalidFlags,
                                ErrUnsupportedType)}


class DeviceFailure(RuntimeError):
    pass


def error_from_code(code: int, position: int = 0):
    if code == 0:
        return None
    if code == 6:
        return ErrOverflow(position)
    if code in _BY_CODE:
        return _BY_CODE[code]
    raise DeviceFailure(f"C-ABI failure {code}: {_lib.last_error()}")


def _buf(src):
    a = np.frombuffer(bytes(src) + b"\0", dtype=np.uint8)
    return a, a.ctypes.data_as(_lib.u8p)


def _h(ctx):
    return ctx.handle if ctx is not None else None


def MaxBlockLength32(flag: int) -> int:
    return _lib.load().utl_max_block_length32(flag)


def MaxBlockLength64(flag: int) -> int:
    return _lib.load().utl_max_block_length64(flag)


def Header(src):
    _a, p = _buf(src)
    out = (C.c_int32 * 7)()
    code = _lib.load().utl_header(p, len(src), out)
    c, bw, exc, d, f, z, s = list(out)
    return c, bw, exc, bool(d), bool(f), bool(z), bool(s), error_from_code(code)


def BlockLength(src):
    _a, p = _buf(src)
    n = C.c_int64()
    code = _lib.load().utl_block_length(p, len(src), C.byref(n))
    return (n.value, None) if code == 0 else (0, error_from_code(code))


def _pack(is64, flag, values, dst, ctx):
    lib = _lib.load()
    prefix = b""
    if flag & Append and dst is not None:
        prefix = bytes(dst)
    arr = np.ascontiguousarray(np.asarray(values, dtype=np.uint64 if is64 else np.uint32))
    work = arr.copy()
    out = np.zeros(2048 if is64 else 1024, dtype=np.uint8)
    n = C.c_size_t()
    ptr = work.ctypes.data_as(_lib.u64p if is64 else _lib.u32p)
    f = lib.utl_pack_u64 if is64 else lib.utl_pack_u32
    code = f(_h(ctx), flag, ptr, work.size, out.ctypes.data_as(_lib.u8p), out.size, C.byref(n))
    if code:
        return None, error_from_code(code)
    if not is64 and not (flag & (NoInPlace | Append)) and isinstance(values, np.ndarray) and \
            values.dtype == np.uint32 and values.flags.writeable:
        values[...] = work  # the reference's in-place result (pack.go:64-66)
    return prefix + out[: n.value].tobytes(), None


def PackUint32(flag: int, values, dst=None, scratch=None, ctx: Context | None = None):
    """pack.go:76. Returns (packed bytes, err). With Append the packed block
    follows dst's bytes. A writable uint32 numpy `values` receives the
    reference's in-place work values unless NoInPlace or Append is set."""
    return _pack(False, flag, values, dst, ctx)


def PackUint64(flag: int, values, dst=None, scratch=None, ctx: Context | None = None):
    """uint64.go:279 (values are never modified)."""
    return _pack(True, flag, values, dst, ctx)


def _unpack(is64, src, values, ctx):
    """Decodes into the caller's array when it can hold the result, as the
    reference does: uint32 when its capacity is at least 128 values
    (growUnpackDst, unpack_block.go:177-182), uint64 when it is at least the
    block's count (uint64.go:224-227); otherwise into a new array. The
    returned values are then a view of the caller's array."""
    lib = _lib.load()
    _a, p = _buf(src)
    dt = np.uint64 if is64 else np.uint32
    need = (src[0] if len(src) else 128) if is64 else 128
    if (isinstance(values, np.ndarray) and values.dtype == dt and values.ndim == 1 and values.flags.c_contiguous
            and values.flags.writeable and values.size >= max(need, 1)):
        dst = values
    else:
        dst = np.zeros(128, dtype=dt)
    c, cons, ovf = C.c_int32(), C.c_int64(), C.c_int64()
    f = lib.utl_unpack_u64 if is64 else lib.utl_unpack_u32
    code = f(_h(ctx), p, len(src), dst.ctypes.data_as(_lib.u64p if is64 else _lib.u32p), C.byref(c),
             C.byref(cons), C.byref(ovf))
    if code:
        return None, 0, error_from_code(code, ovf.value)
    return dst[: c.value], cons.value, None


def UnpackUint32(src, values=None, scratch=None, ctx: Context | None = None):
    """unpack.go:5. Returns (values, consumed, err). A uint32 numpy `values`
    of at least 128 elements receives the decoded values (the result is a
    view of it)."""
    return _unpack(False, src, values, ctx)


def UnpackUint64(src, values=None, scratch=None, ctx: Context | None = None):
    """uint64.go:338. Returns (values, consumed, err). A uint64 numpy `values`
    with room for the block's count receives the decoded values."""
    return _unpack(True, src, values, ctx)


def _get(is64, pos, src, ctx):
    lib = _lib.load()
    _a, p = _buf(src)
    if is64:
        out = C.c_uint64()
        code = lib.utl_get_u64(_h(ctx), pos, p, len(src), C.byref(out))
    else:
        out = C.c_uint32()
        code = lib.utl_get_u32(_h(ctx), pos, p, len(src), C.byref(out))
    return (out.value, None) if code == 0 else (0, error_from_code(code))


def GetUint32(pos: int, src, scratch=None, ctx: Context | None = None):
    """get.go:74. Returns (value, err)."""
    return _get(False, pos, src, ctx)


def GetUint64(pos: int, src, scratch=None, ctx: Context | None = None):
    """uint64.go:364. Returns (value, err)."""
    return _get(True, pos, src, ctx)


class Compressor32:
    """compressor.go:16-56: owns its scratch, here a HIP context (stream and
    staging). Not safe for concurrent use."""

    def __init__(self, device: int = 0):
        self.ctx = Context(device)

    def Compress(self, flags: int, dst, values):
        return PackUint32(flags, values, dst, ctx=self.ctx)

    def Decompress(self, dst, src):
        return UnpackUint32(src, dst, ctx=self.ctx)

    def Get(self, pos: int, src):
        return GetUint32(pos, src, ctx=self.ctx)


class Compressor64:
    """compressor.go:70-107."""

    def __init__(self, device: int = 0):
        self.ctx = Context(device)

    def Compress(self, flags: int, dst, values):
        return PackUint64(flags, values, dst, ctx=self.ctx)

    def Decompress(self, dst, src):
        return UnpackUint64(src, dst, ctx=self.ctx)

    def Get(self, pos: int, src):
        return GetUint64(pos, src, ctx=self.ctx)


def NewUint32(device: int = 0) -> Compressor32:
    return Compressor32(device)


def NewUint64(device: int = 0) -> Compressor64:
    return Compressor64(device)
