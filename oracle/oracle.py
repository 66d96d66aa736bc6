"""ctypes view of the CPU restatement in utl_oracle.c.

TEST INFRASTRUCTURE ONLY: the oracle is the checker the GPU product is
compared against. Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this module.

The functions mirror the Go API shape (file:line of the reference in
utl_oracle.c) and return ``(result..., err)`` with integer error codes.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

OK = 0
ERR_INVALID_BUFFER = 1
ERR_POSITION_OUT_OF_RANGE = 2
ERR_INVALID_BLOCK_LENGTH = 3
ERR_INVALID_FLAGS = 4
ERR_UNSUPPORTED_TYPE = 5
ERR_OVERFLOW = 6
REF_PANIC = 99

Delta, NoFOR, NoPatch, Special, Append, NoInPlace = 1, 2, 4, 8, 16, 32

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


def build() -> str:
    """Compile the restatement with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_header.argtypes = [_u8p, C.c_size_t, _i32p]
        L.orc_block_length.argtypes = [_u8p, C.c_size_t, _i64p]
        L.orc_max_block_length32.argtypes = [C.c_uint32]
        L.orc_max_block_length64.argtypes = [C.c_uint32]
        L.orc_pack_u32.argtypes = [C.c_uint32, _u32p, C.c_size_t, _u8p, C.POINTER(C.c_size_t)]
        L.orc_unpack_u32.argtypes = [_u8p, C.c_size_t, _u32p, _i32p, _i64p, _i64p]
        L.orc_get_u32.argtypes = [C.c_int64, _u8p, C.c_size_t, _u32p]
        L.orc_pack_u64.argtypes = [C.c_uint32, _u64p, C.c_size_t, _u8p, C.POINTER(C.c_size_t)]
        L.orc_unpack_u64.argtypes = [_u8p, C.c_size_t, _u64p, _i32p, _i64p, _i64p]
        L.orc_get_u64.argtypes = [C.c_int64, _u8p, C.c_size_t, _u64p]
        for f in (L.orc_pack_stream_u32, L.orc_pack_stream_u64):
            f.argtypes = [C.c_uint32, C.c_void_p, C.c_size_t, _u8p, C.c_size_t, _u32p, _u32p, C.c_int]
            f.restype = C.c_int64
        for f in (L.orc_unpack_blocks_u32, L.orc_unpack_blocks_u64):
            f.argtypes = [_u8p, _u64p, _u64p, C.c_size_t, C.c_void_p, _i32p, C.c_int]
            f.restype = None
        L.orc_walk_stream.argtypes = [_u8p, C.c_size_t, _u64p, _u32p, _u32p, C.c_size_t, _i32p]
        L.orc_walk_stream.restype = C.c_int64
        _lib = L
    return _lib


def _bytes_ptr(b):
    a = np.frombuffer(bytes(b) + b"\0", dtype=np.uint8)  # +1 keeps empty buffers addressable
    return a, a.ctypes.data_as(_u8p)


def Header(src: bytes):
    _a, p = _bytes_ptr(src)
    out = (C.c_int32 * 7)()
    err = lib().orc_header(p, len(src), out)
    c, bw, exc, d, f, z, s = list(out)
    return c, bw, exc, bool(d), bool(f), bool(z), bool(s), err


def BlockLength(src: bytes):
    _a, p = _bytes_ptr(src)
    n = C.c_int64()
    err = lib().orc_block_length(p, len(src), C.byref(n))
    return n.value, err


def MaxBlockLength32(flag: int) -> int:
    return lib().orc_max_block_length32(flag)


def MaxBlockLength64(flag: int) -> int:
    return lib().orc_max_block_length64(flag)


def PackUint32(flag: int, values):
    """Returns (packed bytes, work values after the call, err)."""
    v = np.ascontiguousarray(np.array(values, dtype=np.uint32).copy())
    dst = np.zeros(2048, dtype=np.uint8)
    n = C.c_size_t()
    vp = v.ctypes.data_as(_u32p) if v.size else C.cast(dst.ctypes.data, _u32p)
    err = lib().orc_pack_u32(flag, vp, v.size, dst.ctypes.data_as(_u8p), C.byref(n))
    if err:
        return None, v, err
    return dst[: n.value].tobytes(), v, err


def UnpackUint32(src: bytes):
    """Returns (values, consumed, err, overflow_pos)."""
    _a, p = _bytes_ptr(src)
    dst = np.zeros(128, dtype=np.uint32)
    c, cons, ovf = C.c_int32(), C.c_int64(), C.c_int64()
    err = lib().orc_unpack_u32(p, len(src), dst.ctypes.data_as(_u32p), C.byref(c), C.byref(cons), C.byref(ovf))
    if err:
        return None, 0, err, ovf.value
    return dst[: c.value].copy(), cons.value, err, 0


def GetUint32(pos: int, src: bytes):
    _a, p = _bytes_ptr(src)
    out = C.c_uint32()
    err = lib().orc_get_u32(pos, p, len(src), C.byref(out))
    return out.value, err


def PackUint64(flag: int, values):
    v = np.ascontiguousarray(np.array(values, dtype=np.uint64))
    dst = np.zeros(4096, dtype=np.uint8)
    n = C.c_size_t()
    vp = v.ctypes.data_as(_u64p) if v.size else C.cast(dst.ctypes.data, _u64p)
    err = lib().orc_pack_u64(flag, vp, v.size, dst.ctypes.data_as(_u8p), C.byref(n))
    if err:
        return None, err
    return dst[: n.value].tobytes(), err


def UnpackUint64(src: bytes):
    _a, p = _bytes_ptr(src)
    dst = np.zeros(128, dtype=np.uint64)
    c, cons, ovf = C.c_int32(), C.c_int64(), C.c_int64()
    err = lib().orc_unpack_u64(p, len(src), dst.ctypes.data_as(_u64p), C.byref(c), C.byref(cons), C.byref(ovf))
    if err:
        return None, 0, err, ovf.value
    return dst[: c.value].copy(), cons.value, err, 0


def GetUint64(pos: int, src: bytes):
    _a, p = _bytes_ptr(src)
    out = C.c_uint64()
    err = lib().orc_get_u64(pos, p, len(src), C.byref(out))
    return out.value, err


def pack_stream(flag: int, values: np.ndarray, nthreads: int = 8):
    """Concatenate 128-value blocks (flag|Append). Returns
    (stream bytes as uint8 array, block_len uint32, block_count uint32)."""
    is64 = values.dtype == np.uint64
    values = np.ascontiguousarray(values)
    n = values.size
    nb = (n + 127) // 128
    cap = nb * (2038 if is64 else 1018) + 16
    dst = np.zeros(cap, dtype=np.uint8)
    bl = np.zeros(nb, dtype=np.uint32)
    bc = np.zeros(nb, dtype=np.uint32)
    f = lib().orc_pack_stream_u64 if is64 else lib().orc_pack_stream_u32
    r = f(flag, values.ctypes.data, n, dst.ctypes.data_as(_u8p), cap, bl.ctypes.data_as(_u32p),
          bc.ctypes.data_as(_u32p), nthreads)
    if r < 0:
        raise RuntimeError(f"oracle pack_stream failed: {-r}")
    return dst[:r], bl, bc


def unpack_blocks(src: np.ndarray, off: np.ndarray, out_off: np.ndarray, n_values: int, is64: bool,
                  nthreads: int = 8):
    """Decode block k = Unpack(src[off[k]:off[k+1]]) into out[out_off[k]:]."""
    nb = off.size - 1
    out = np.zeros(n_values, dtype=np.uint64 if is64 else np.uint32)
    status = np.zeros(nb, dtype=np.int32)
    f = lib().orc_unpack_blocks_u64 if is64 else lib().orc_unpack_blocks_u32
    src = np.ascontiguousarray(src, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    out_off = np.ascontiguousarray(out_off, dtype=np.uint64)
    f(src.ctypes.data_as(_u8p), off.ctypes.data_as(_u64p), out_off.ctypes.data_as(_u64p), nb,
      out.ctypes.data, status.ctypes.data_as(_i32p), nthreads)
    return out, status


def walk_stream(src: np.ndarray):
    """Serial BlockLength walk of one stream (example_test.go:171-187).
    Returns (starts uint64, lengths uint32, counts uint32, status)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    pad = np.concatenate([src, np.zeros(1, np.uint8)])  # keeps empty streams addressable
    cap = src.size // 4 + 1  # a block is at least 4 bytes
    starts = np.zeros(cap, np.uint64)
    lens = np.zeros(cap, np.uint32)
    counts = np.zeros(cap, np.uint32)
    st = C.c_int32()
    n = lib().orc_walk_stream(pad.ctypes.data_as(_u8p), src.size, starts.ctypes.data_as(_u64p),
                              lens.ctypes.data_as(_u32p), counts.ctypes.data_as(_u32p), cap, C.byref(st))
    return starts[:n], lens[:n], counts[:n], st.value
