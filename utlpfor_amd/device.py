"""Device memory and context handles over the C-ABI (no torch: the library
owns its HIP runtime, so one process never mixes two HIP runtimes)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class UtlRuntimeError(RuntimeError):
    pass


def check(code: int, what: str) -> None:
    if code != 0:
        raise UtlRuntimeError(f"{what} failed with code {code}: {_lib.last_error()}")


class Context:
    """One HIP stream plus staging memory (a Go Compressor's scratch)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.device = device
        self.handle = self.lib.utl_ctx_create(device)
        if not self.handle:
            raise UtlRuntimeError(f"no usable gfx950 device {device}: {_lib.last_error()}")

    def sync(self) -> None:
        check(self.lib.utl_ctx_sync(self.handle), "utl_ctx_sync")

    @property
    def stream(self) -> int:
        return self.lib.utl_ctx_stream(self.handle)

    def device_info(self):
        name = C.create_string_buffer(128)
        cu = C.c_int()
        check(self.lib.utl_ctx_device_info(self.handle, name, 128, C.byref(cu)), "utl_ctx_device_info")
        return name.value.decode(), cu.value

    def close(self) -> None:
        if self.handle:
            self.lib.utl_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation with a numpy dtype and shape (1-D)."""

    def __init__(self, n: int, dtype, ctx: Context | None = None):
        self.lib = _lib.load()
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        self.ctx = ctx
        p = C.c_void_p()
        check(self.lib.utl_dev_alloc(C.byref(p), max(self.nbytes, 1)), "utl_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray, ctx: Context | None = None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, ctx)
        if a.nbytes:
            check(d.lib.utl_memcpy_h2d(ctx.handle if ctx else None, d.ptr, a.ctypes.data, a.nbytes), "h2d")
        return d

    def to_numpy(self, count: int | None = None) -> np.ndarray:
        n = self.n if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if out.nbytes:
            check(self.lib.utl_memcpy_d2h(self.ctx.handle if self.ctx else None, out.ctypes.data, self.ptr,
                                          out.nbytes), "d2h")
        return out

    def read(self, first: int, count: int) -> np.ndarray:
        """Elements [first, first + count) as a host array."""
        if first < 0 or count < 0 or first + count > self.n:
            raise IndexError(f"read [{first}, {first + count}) of {self.n}")
        out = np.empty(count, dtype=self.dtype)
        if out.nbytes:
            check(self.lib.utl_memcpy_d2h(self.ctx.handle if self.ctx else None, out.ctypes.data,
                                          self.ptr + first * self.dtype.itemsize, out.nbytes), "d2h")
        return out

    def write(self, first: int, a: np.ndarray) -> None:
        """Copies host array `a` into elements [first, first + a.size)."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if first < 0 or first + a.size > self.n:
            raise IndexError(f"write [{first}, {first + a.size}) of {self.n}")
        if a.nbytes:
            check(self.lib.utl_memcpy_h2d(self.ctx.handle if self.ctx else None,
                                          self.ptr + first * self.dtype.itemsize, a.ctypes.data, a.nbytes), "h2d")

    def fill_bytes(self, value: int) -> None:
        check(self.lib.utl_memset_d(self.ctx.handle if self.ctx else None, self.ptr, value, self.nbytes), "fill")

    def free(self) -> None:
        if self.ptr:
            self.lib.utl_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    """A HIP event recorded on a context's stream."""

    def __init__(self):
        self.lib = _lib.load()
        p = C.c_void_p()
        check(self.lib.utl_event_create(C.byref(p)), "utl_event_create")
        self.ptr = p.value

    def record(self, ctx: Context) -> None:
        check(self.lib.utl_event_record(ctx.handle, self.ptr), "utl_event_record")

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        check(self.lib.utl_event_elapsed_ms(self.ptr, stop.ptr, C.byref(ms)), "utl_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            self.lib.utl_event_destroy(self.ptr)
        except Exception:
            pass


def dev_copy(ctx: Context, dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
    check(_lib.load().utl_dev_copy(ctx.handle, dst.ptr, src.ptr, nbytes), "utl_dev_copy")


def dev_stream(ctx: Context, dst: DeviceArray, write_bytes: int, src: DeviceArray, read_bytes: int,
               grid_wgs: int = 0) -> None:
    """The streaming kernel of a given traffic shape (utl_dev_stream)."""
    check(_lib.load().utl_dev_stream(ctx.handle, dst.ptr, write_bytes, src.ptr, read_bytes, grid_wgs),
          "utl_dev_stream")


class PinnedBuffer:
    """Page-locked host memory (utl_host_alloc): device copies into it run
    at the full link rate (on MI355X about 57 GB/s against 10-17 GB/s into
    pageable memory)."""

    def __init__(self, nbytes: int):
        self.lib = _lib.load()
        p = C.c_void_p()
        check(self.lib.utl_host_alloc(C.byref(p), max(int(nbytes), 1)), "utl_host_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.bytes = np.ctypeslib.as_array((C.c_uint8 * max(self.nbytes, 1)).from_address(self.ptr))

    def free(self) -> None:
        if self.ptr:
            self.bytes = None
            self.lib.utl_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mem_info():
    lib = _lib.load()
    f, t = C.c_size_t(), C.c_size_t()
    check(lib.utl_dev_mem_info(C.byref(f), C.byref(t)), "utl_dev_mem_info")
    return f.value, t.value
