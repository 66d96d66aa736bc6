"""ctypes binding of libutlpfor_gpu.so (the C-ABI of include/utlpfor_gpu.h).

The library is loaded from this package directory only. There is no CPU
fallback: if the library is missing, or no gfx950 device is present when a
compute entry point runs, the call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libutlpfor_gpu.so")

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
vp = C.c_void_p
sz = C.c_size_t

# name: (restype, argtypes)
SIGNATURES = {
    "utl_ctx_create": (vp, [C.c_int]),
    "utl_ctx_destroy": (None, [vp]),
    "utl_ctx_stream": (vp, [vp]),
    "utl_ctx_sync": (C.c_int, [vp]),
    "utl_ctx_device_info": (C.c_int, [vp, C.c_char_p, sz, C.POINTER(C.c_int)]),
    "utl_max_block_length32": (C.c_int, [C.c_uint32]),
    "utl_max_block_length64": (C.c_int, [C.c_uint32]),
    "utl_header": (C.c_int, [u8p, sz, i32p]),
    "utl_block_length": (C.c_int, [u8p, sz, i64p]),
    "utl_pack_u32": (C.c_int, [vp, C.c_uint32, u32p, sz, u8p, sz, C.POINTER(sz)]),
    "utl_unpack_u32": (C.c_int, [vp, u8p, sz, u32p, i32p, i64p, i64p]),
    "utl_get_u32": (C.c_int, [vp, C.c_int64, u8p, sz, u32p]),
    "utl_pack_u64": (C.c_int, [vp, C.c_uint32, u64p, sz, u8p, sz, C.POINTER(sz)]),
    "utl_unpack_u64": (C.c_int, [vp, u8p, sz, u64p, i32p, i64p, i64p]),
    "utl_get_u64": (C.c_int, [vp, C.c_int64, u8p, sz, u64p]),
    "utl_dev_alloc": (C.c_int, [C.POINTER(vp), sz]),
    "utl_dev_free": (C.c_int, [vp]),
    "utl_host_alloc": (C.c_int, [C.POINTER(vp), sz]),
    "utl_host_free": (C.c_int, [vp]),
    "utl_memcpy_h2d": (C.c_int, [vp, vp, vp, sz]),
    "utl_memcpy_d2h": (C.c_int, [vp, vp, vp, sz]),
    "utl_memset_d": (C.c_int, [vp, vp, C.c_int, sz]),
    "utl_dev_mem_info": (C.c_int, [C.POINTER(sz), C.POINTER(sz)]),
    "utl_index_build": (C.c_int, [vp, vp, sz, vp, vp]),
    "utl_block_lengths": (C.c_int, [vp, vp, vp, vp, sz, vp, vp]),
    "utl_decode_u32": (C.c_int, [vp, vp, vp, vp, sz, vp, vp, vp]),
    "utl_decode_u64": (C.c_int, [vp, vp, vp, vp, sz, vp, vp, vp]),
    "utl_decode_persistent": (C.c_int, [vp, C.c_int, vp, vp, vp, sz, vp, vp, vp]),
    "utl_wave_index_build": (C.c_int, [vp, vp, sz, vp]),
    "utl_index_build_noalloc": (C.c_int, [vp, vp, sz, vp, vp]),
    "utl_decode_meta": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, sz, vp, vp, vp]),
    "utl_set_tune": (C.c_int, [vp, vp]),
    "utl_get_tune": (C.c_int, [vp, vp]),
    "utl_get_batch": (C.c_int, [vp, C.c_int, vp, vp, vp, sz, vp, vp, sz, vp, vp]),
    "utl_pos_dir_build": (C.c_int, [vp, vp, vp, sz, vp, sz]),
    "utl_get_batch_dir": (C.c_int, [vp, C.c_int, vp, vp, vp, sz, vp, sz, vp, sz, vp, vp]),
    "utl_gather_index": (C.c_int, [vp, vp, vp, sz, vp, sz, sz, vp, vp]),
    "utl_index_walk": (C.c_int, [vp, vp, vp, sz, vp, vp, vp, vp, vp]),
    "utl_encode_blocks": (C.c_int, [vp, C.c_int, vp, vp, vp, sz, vp, vp]),
    "utl_compact": (C.c_int, [vp, C.c_int, vp, vp, sz, vp]),
    "utl_encode_u32_size": (C.c_int, [vp, vp, vp, vp, sz, vp, vp]),
    "utl_encode_u32_write": (C.c_int, [vp, vp, vp, vp, sz, vp, vp, vp, sz]),
    "utl_encode_u64_size": (C.c_int, [vp, vp, vp, vp, sz, vp]),
    "utl_encode_u64_write": (C.c_int, [vp, vp, vp, vp, sz, vp, vp, sz]),
    "utl_encode_u32_direct": (C.c_int, [vp, vp, vp, vp, sz, vp, sz, vp, vp, C.POINTER(sz)]),
    "utl_make_jobs": (C.c_int, [vp, vp, vp, vp, sz, vp, vp]),
    "utl_graph_decode_create": (vp, [vp, C.c_int, C.c_int, vp, vp, sz, vp, vp, vp, vp]),
    "utl_graph_launch": (C.c_int, [vp]),
    "utl_graph_destroy": (None, [vp]),
    "utl_event_create": (C.c_int, [C.POINTER(vp)]),
    "utl_event_destroy": (C.c_int, [vp]),
    "utl_event_record": (C.c_int, [vp, vp]),
    "utl_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "utl_dev_copy": (C.c_int, [vp, vp, vp, sz]),
    "utl_dev_stream": (C.c_int, [vp, vp, sz, vp, sz, C.c_int]),
    "utl_container_bytes": (sz, [sz, sz, sz]),
    "utl_container_write": (C.c_int, [vp, sz, vp, sz, vp, sz, vp, sz, C.POINTER(sz)]),
    "utl_container_parse": (C.c_int, [vp, sz, vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "utl_last_error": (C.c_char_p, []),
}

_lib = None


def load() -> C.CDLL:
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m utlpfor_amd.build` "
                               "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return load().utl_last_error().decode(errors="replace")
