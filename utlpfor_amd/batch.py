"""Device-resident streams: batched encode, index, decode and Get.

A DeviceStream is the concatenation of UTL-PFOR blocks exactly as a caller
of the reference builds it with PackUint32(flag|Append, ...) per 128-value
chunk (pack.go:53-76), kept in HBM together with its index: per-block meta
words (byte length | count << 24, produced by the encoder), block byte
offsets and decoded-value offsets (the device scan that replaces the serial
BlockLength walk of example_test.go:171-187).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import Context, DeviceArray, check

SLOT32, SLOT64 = 1024, 2048
MAX_BLOCK32 = 1018  # MaxBlockLength32(0), the longest uint32 block (scalar.go:336-357)
MAX_BLOCK64 = 2038  # MaxBlockLength64(0), the longest uint64 block (uint64.go:451-453)
POS_DIR_WORDS = 4  # uint64 words per position-directory entry (UTL_POS_DIR_WORDS)


@dataclass
class DeviceStream:
    ctx: Context
    data: DeviceArray       # uint8, concatenated blocks
    meta: DeviceArray       # uint32 per block
    block_off: DeviceArray  # uint64, nblocks + 1
    out_off: DeviceArray    # uint64, nblocks + 1
    nblocks: int
    nvalues: int
    nbytes: int
    is64: bool
    pos_dir: DeviceArray | None = None  # position directory, built by the first Get by stream position

    def bytes_host(self) -> np.ndarray:
        return self.data.to_numpy(self.nbytes)

    def ensure_pos_dir(self) -> DeviceArray:
        """The position directory over block_off / out_off (utl_pos_dir_build):
        entry c (POS_DIR_WORDS words) describes the block holding value
        position 128c and the block after it."""
        if self.pos_dir is None:
            lib = _lib.load()
            ndir = max((self.nvalues + 127) // 128, 1)
            d = DeviceArray(ndir * POS_DIR_WORDS, np.uint64, self.ctx)
            check(lib.utl_pos_dir_build(self.ctx.handle, self.block_off.ptr, self.out_off.ptr, self.nblocks, d.ptr,
                                        ndir), "utl_pos_dir_build")
            self.pos_dir = d
        return self.pos_dir

    def free(self) -> None:
        for a in (self.data, self.meta, self.block_off, self.out_off, self.pos_dir):
            if a is not None:
                a.free()
        self.pos_dir = None


def plan_blocks(stream_lengths: np.ndarray) -> np.ndarray:
    """Block numbering of streams cut into 128-value chunks: blk_off[s]."""
    nb = (np.asarray(stream_lengths, dtype=np.uint64) + 127) // 128
    return np.concatenate([[0], np.cumsum(nb)]).astype(np.uint64)


def build_index(ctx: Context, meta: DeviceArray, nblocks: int):
    lib = _lib.load()
    boff = DeviceArray(nblocks + 1, np.uint64, ctx)
    ooff = DeviceArray(nblocks + 1, np.uint64, ctx)
    check(lib.utl_index_build(ctx.handle, meta.ptr, nblocks, boff.ptr, ooff.ptr), "utl_index_build")
    return boff, ooff


ENCODE_METHODS = ("slot", "direct", "sized")


def encode_method(ctx: Context, is64: bool, nblocks: int, nvalues: int, method: str = "auto") -> str:
    """The encode form to use: "slot" (each block into a fixed slot, then the
    index scan and a compaction pass), "direct" (uint32 only: blocks
    written straight to their place, offsets by look-back) or "sized"
    (a size pass, the index scan, a write pass at the offsets). "auto"
    takes the autotuner's cached choice for this shape of uint32 work, else
    the sized form."""
    if method in ENCODE_METHODS:
        return "slot" if is64 and method == "direct" else method
    from . import autotune
    return (None if is64 else autotune.cached_encode_method(ctx, nblocks, nvalues)) or "sized"


def encode(ctx: Context, values, flag: int = 0, stream_lengths=None, stream_flags=None,
           is64: bool | None = None, method: str = "auto") -> DeviceStream:
    """Encodes one stream (or several, back to back) on the GPU. `values` is a
    numpy array or a DeviceArray of uint32/uint64. `method`: see
    encode_method; every form gives identical bytes."""
    lib = _lib.load()
    dv = values if isinstance(values, DeviceArray) else DeviceArray.from_numpy(np.ascontiguousarray(values), ctx)
    if is64 is None:
        is64 = dv.dtype == np.uint64
    n = dv.n
    if stream_lengths is None:
        stream_lengths = np.array([n], dtype=np.uint64)
    stream_lengths = np.asarray(stream_lengths, dtype=np.uint64)
    if stream_flags is None:
        stream_flags = np.full(stream_lengths.size, flag, dtype=np.uint8)
    blk_off = plan_blocks(stream_lengths)
    nblocks = int(blk_off[-1])
    val_off = np.concatenate([[0], np.cumsum(stream_lengths)]).astype(np.uint64)
    d_val_off = DeviceArray.from_numpy(val_off, ctx)
    d_blk_off = DeviceArray.from_numpy(blk_off, ctx)
    d_flags = DeviceArray.from_numpy(np.asarray(stream_flags, dtype=np.uint8), ctx)
    in_off = DeviceArray(nblocks, np.uint64, ctx)
    in_meta = DeviceArray(nblocks, np.uint32, ctx)
    check(lib.utl_make_jobs(ctx.handle, d_val_off.ptr, d_blk_off.ptr, d_flags.ptr, stream_lengths.size, in_off.ptr,
                            in_meta.ptr), "utl_make_jobs")
    meta = DeviceArray(nblocks, np.uint32, ctx)
    form = encode_method(ctx, bool(is64), nblocks, n, method)
    if form == "sized" and is64:
        # lengths, the index scan, then every block encoded again at its place
        check(lib.utl_encode_u64_size(ctx.handle, dv.ptr, in_off.ptr, in_meta.ptr, nblocks, meta.ptr),
              "utl_encode_u64_size")
        boff, ooff = build_index(ctx, meta, nblocks)
        total = int(boff.read(nblocks, 1)[0]) if nblocks else 0
        data = DeviceArray(total + 64, np.uint8, ctx)
        check(lib.utl_encode_u64_write(ctx.handle, dv.ptr, in_off.ptr, in_meta.ptr, nblocks, boff.ptr, data.ptr,
                                       data.nbytes), "utl_encode_u64_write")
        scratch = ()
    elif form == "sized":
        # decisions and lengths, the index scan, then every block at its place
        plan = DeviceArray(max(nblocks, 1), np.uint64, ctx)
        check(lib.utl_encode_u32_size(ctx.handle, dv.ptr, in_off.ptr, in_meta.ptr, nblocks, meta.ptr, plan.ptr),
              "utl_encode_u32_size")
        boff, ooff = build_index(ctx, meta, nblocks)
        total = int(boff.read(nblocks, 1)[0]) if nblocks else 0
        data = DeviceArray(total + 64, np.uint8, ctx)
        check(lib.utl_encode_u32_write(ctx.handle, dv.ptr, in_off.ptr, in_meta.ptr, nblocks, plan.ptr, boff.ptr,
                                       data.ptr, data.nbytes), "utl_encode_u32_write")
        scratch = (plan,)
    elif form == "direct":
        # the stream written in place; the index scan then gives both offsets
        data = DeviceArray(nblocks * MAX_BLOCK32 + 64, np.uint8, ctx)
        dboff = DeviceArray(nblocks + 1, np.uint64, ctx)
        nb_out = C.c_size_t()
        check(lib.utl_encode_u32_direct(ctx.handle, dv.ptr, in_off.ptr, in_meta.ptr, nblocks, data.ptr, data.nbytes,
                                        meta.ptr, dboff.ptr, C.byref(nb_out)), "utl_encode_u32_direct")
        dboff.free()
        boff, ooff = build_index(ctx, meta, nblocks)
        total = nb_out.value
        scratch = ()
    else:
        slot = SLOT64 if is64 else SLOT32
        slots = DeviceArray(nblocks * slot, np.uint8, ctx)
        check(lib.utl_encode_blocks(ctx.handle, int(is64), dv.ptr, in_off.ptr, in_meta.ptr, nblocks, slots.ptr,
                                    meta.ptr), "utl_encode_blocks")
        boff, ooff = build_index(ctx, meta, nblocks)
        total = int(boff.read(nblocks, 1)[0]) if nblocks else 0
        data = DeviceArray(total + 64, np.uint8, ctx)
        check(lib.utl_compact(ctx.handle, int(is64), slots.ptr, boff.ptr, nblocks, data.ptr), "utl_compact")
        scratch = (slots,)
    ctx.sync()
    for a in (*scratch, in_off, in_meta, d_val_off, d_blk_off, d_flags):
        a.free()
    if values is not dv:
        dv.free()
    return DeviceStream(ctx, data, meta, boff, ooff, nblocks, n, total, bool(is64))


def from_host_blocks(ctx: Context, stream_bytes: np.ndarray, block_len: np.ndarray, block_count: np.ndarray,
                     is64: bool = False) -> DeviceStream:
    """Uploads an already-encoded stream with its per-block lengths/counts."""
    nblocks = int(block_len.size)
    meta_h = (block_len.astype(np.uint32) & 0xFFFFFF) | (block_count.astype(np.uint32) << 24)
    meta = DeviceArray.from_numpy(meta_h, ctx)
    boff, ooff = build_index(ctx, meta, nblocks)
    data = DeviceArray.from_numpy(np.concatenate([np.asarray(stream_bytes, dtype=np.uint8),
                                                  np.zeros(64, np.uint8)]), ctx)
    ctx.sync()
    return DeviceStream(ctx, data, meta, boff, ooff, nblocks, int(block_count.sum()), int(stream_bytes.size),
                        is64)


def index_foreign(ctx: Context, data: DeviceArray, stream_off: np.ndarray, is64: bool = False):
    """Indexes streams of concatenated blocks whose block starts are unknown
    (stream s = data[stream_off[s]:stream_off[s+1]]) with the device walk.
    Returns (DeviceStream over all walked blocks, per-stream block counts,
    per-stream walk statuses). A stream's blocks after its stop (its first
    BlockLength error or a truncated block) are not indexed. block_off[k] is
    the walked start of block k, so block k decodes from its start to the
    next walked block's start (its exact end inside a stream, the rest of the
    stream after a stop, like UnpackUint32(buf[pos:]) in example_test.go:183);
    the last entry is the last block's end. `data` must be 16-byte aligned."""
    lib = _lib.load()
    stream_off = np.ascontiguousarray(stream_off, dtype=np.uint64)
    ns = stream_off.size - 1
    d_off = DeviceArray.from_numpy(stream_off, ctx)
    d_n = DeviceArray(max(ns, 1), np.uint32, ctx)
    d_st = DeviceArray(max(ns, 1), np.int32, ctx)
    check(lib.utl_index_walk(ctx.handle, data.ptr, d_off.ptr, ns, None, None, None, d_n.ptr, d_st.ptr),
          "walk (count)")
    counts = d_n.to_numpy(ns).astype(np.uint64)
    status = d_st.to_numpy(ns)
    blk_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    nblocks = int(blk_off[-1])
    d_blk = DeviceArray.from_numpy(blk_off, ctx)
    meta = DeviceArray(max(nblocks, 1), np.uint32, ctx)
    boff = DeviceArray(nblocks + 1, np.uint64, ctx)
    check(lib.utl_index_walk(ctx.handle, data.ptr, d_off.ptr, ns, d_blk.ptr, meta.ptr, boff.ptr, d_n.ptr,
                             d_st.ptr), "walk (meta)")
    scanned, ooff = build_index(ctx, meta, nblocks)
    nbytes = 0
    if nblocks:
        last_start = int(boff.read(nblocks - 1, 1)[0])
        last_len = int(meta.read(nblocks - 1, 1)[0]) & 0xFFFFFF
        nbytes = last_start + last_len
    boff.write(nblocks, np.array([nbytes], dtype=np.uint64))
    nvalues = int(ooff.read(nblocks, 1)[0])
    ctx.sync()
    for a in (d_off, d_n, d_st, d_blk, scanned):
        a.free()
    return DeviceStream(ctx, data, meta, boff, ooff, nblocks, nvalues, nbytes, is64), counts, status


def decode(ctx: Context, s: DeviceStream, out: DeviceArray | None = None, status: DeviceArray | None = None,
           consumed: DeviceArray | None = None, persistent: bool = False):
    """One launch decoding every block of the stream into `out`."""
    lib = _lib.load()
    if out is None:
        out = DeviceArray(max(s.nvalues, 1), np.uint64 if s.is64 else np.uint32, ctx)
    if status is None:
        status = DeviceArray(max(s.nblocks, 1), np.int32, ctx)
    cptr = consumed.ptr if consumed is not None else None
    if persistent:
        code = lib.utl_decode_persistent(ctx.handle, int(s.is64), s.data.ptr, s.block_off.ptr, s.out_off.ptr,
                                         s.nblocks, out.ptr, status.ptr, cptr)
    elif s.is64:
        code = lib.utl_decode_u64(ctx.handle, s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks, out.ptr,
                                  status.ptr, cptr)
    else:
        code = lib.utl_decode_u32(ctx.handle, s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks, out.ptr,
                                  status.ptr, cptr)
    check(code, "decode")
    return out, status


def build_wave_index(ctx: Context, s: DeviceStream, wave: DeviceArray | None = None) -> DeviceArray:
    """The wave index of s (byte and value offsets of every eighth block)."""
    if wave is None:
        wave = DeviceArray(2 * ((s.nblocks + 7) // 8) + 2, np.uint64, ctx)
    check(_lib.load().utl_wave_index_build(ctx.handle, s.meta.ptr, s.nblocks, wave.ptr), "utl_wave_index_build")
    return wave


def decode_meta(ctx: Context, s: DeviceStream, wave: DeviceArray, out: DeviceArray | None = None,
                status: DeviceArray | None = None, consumed: DeviceArray | None = None, persistent: bool = False):
    """One launch decoding every block of s from its meta words and wave index."""
    if out is None:
        out = DeviceArray(max(s.nvalues, 1), np.uint64 if s.is64 else np.uint32, ctx)
    if status is None:
        status = DeviceArray(max(s.nblocks, 1), np.int32, ctx)
    cptr = consumed.ptr if consumed is not None else None
    check(_lib.load().utl_decode_meta(ctx.handle, int(s.is64), int(persistent), s.data.ptr, s.meta.ptr, wave.ptr,
                                      s.nblocks, out.ptr, status.ptr, cptr), "utl_decode_meta")
    return out, status


def get_batch(ctx: Context, s: DeviceStream, positions: DeviceArray, out: DeviceArray | None = None,
              status: DeviceArray | None = None, blocks: DeviceArray | None = None, use_dir: bool | None = None):
    """Batched Get of stream positions (or of (block, in-block pos) pairs).
    Stream positions go through the stream's position directory
    (utl_get_batch_dir) when its block numbers drift from position / 128
    (more than 16 blocks beyond ceil(values / 128): partial blocks inside
    the stream, as in a store of many streams), else through utl_get_batch's
    search from position / 128, which is exact for a stream of full blocks
    and saves the directory read. use_dir True / False forces either."""
    lib = _lib.load()
    n = positions.n
    if out is None:
        out = DeviceArray(max(n, 1), np.uint64 if s.is64 else np.uint32, ctx)
    if status is None:
        status = DeviceArray(max(n, 1), np.int32, ctx)
    if use_dir is None:
        use_dir = s.nblocks > (s.nvalues + 127) // 128 + 16
    if blocks is None and use_dir and s.nblocks > 0:
        d = s.ensure_pos_dir()
        check(lib.utl_get_batch_dir(ctx.handle, int(s.is64), s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks,
                                    d.ptr, d.n // POS_DIR_WORDS, positions.ptr, n, out.ptr, status.ptr),
              "utl_get_batch_dir")
        return out, status
    check(lib.utl_get_batch(ctx.handle, int(s.is64), s.data.ptr, s.block_off.ptr, s.out_off.ptr, s.nblocks,
                            blocks.ptr if blocks is not None else None, positions.ptr, n, out.ptr, status.ptr),
          "utl_get_batch")
    return out, status


class DecodeGraph:
    """The whole decode of a resident stream (device index scan + decode)
    captured once as a HIP graph and replayed with one launch. With
    block_index False the graph keeps only the wave index (one entry per
    eight blocks) and decodes from the meta words; otherwise it rewrites the
    stream's block_off / out_off and decodes from them."""

    def __init__(self, ctx: Context, s: DeviceStream, out: DeviceArray, status: DeviceArray,
                 persistent: bool = True, block_index: bool = True):
        lib = _lib.load()
        self.ctx = ctx
        bo = s.block_off.ptr if block_index else None
        oo = s.out_off.ptr if block_index else None
        self.handle = lib.utl_graph_decode_create(ctx.handle, int(s.is64), int(persistent), s.data.ptr, s.meta.ptr,
                                                  s.nblocks, bo, oo, out.ptr, status.ptr)
        if not self.handle:
            raise RuntimeError(f"graph capture failed: {_lib.last_error()}")

    def launch(self) -> None:
        check(_lib.load().utl_graph_launch(self.handle), "utl_graph_launch")

    def close(self) -> None:
        if self.handle:
            _lib.load().utl_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TuneParams(C.Structure):
    """utl_tune_params of the C-ABI: the decode launch variant."""
    _fields_ = [("arena_bytes", C.c_int), ("waves_per_wg", C.c_int), ("grid_wgs", C.c_int),
                ("tiles_per_grab", C.c_int)]


ARENAS32 = (2560, 3328, 4352)
WAVES_PER_WG = (2, 4)


def set_tune(ctx: Context, arena_bytes: int, waves_per_wg: int, grid_wgs: int, tiles_per_grab: int) -> None:
    t = TuneParams(arena_bytes, waves_per_wg, grid_wgs, tiles_per_grab)
    check(_lib.load().utl_set_tune(ctx.handle, C.byref(t)), "utl_set_tune")


def get_tune(ctx: Context) -> dict:
    t = TuneParams()
    check(_lib.load().utl_get_tune(ctx.handle, C.byref(t)), "utl_get_tune")
    return {"arena_bytes": t.arena_bytes, "waves_per_wg": t.waves_per_wg, "grid_wgs": t.grid_wgs,
            "tiles_per_grab": t.tiles_per_grab}


# ---- stream container (utl_container_* in the C-ABI) ----

STREAM_ENTRY = np.dtype([("first_block", "<u8"), ("nblocks", "<u8"), ("first_value", "<u8"), ("nvalues", "<u8"),
                         ("byte_off", "<u8"), ("flag", "<u4"), ("is_u64", "<u4")])
CONTAINER_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"), ("nstreams", "<u8"),
                             ("nblocks", "<u8"), ("nbytes", "<u8"), ("nvalues", "<u8"), ("dir_off", "<u8"),
                             ("meta_off", "<u8"), ("data_off", "<u8"), ("total_bytes", "<u8")])
assert STREAM_ENTRY.itemsize == 48 and CONTAINER_HEADER.itemsize == 80


def directory(s: DeviceStream, stream_lengths, stream_flags=None) -> np.ndarray:
    """Stream directory of a stream set encoded by encode() with these
    per-stream value counts (and flags)."""
    lens = np.asarray(stream_lengths, dtype=np.uint64)
    blk = plan_blocks(lens)
    boff = s.block_off.to_numpy()
    d = np.zeros(lens.size, STREAM_ENTRY)
    d["first_block"] = blk[:-1]
    d["nblocks"] = np.diff(blk)
    d["first_value"] = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64)
    d["nvalues"] = lens
    d["byte_off"] = boff[blk[:-1].astype(np.int64)]
    d["flag"] = 0 if stream_flags is None else np.asarray(stream_flags, dtype=np.uint32)
    d["is_u64"] = int(s.is64)
    return d


def to_container(s: DeviceStream, dir_: np.ndarray) -> np.ndarray:
    """The container bytes (uint8 array) of a device stream set."""
    lib = _lib.load()
    dir_ = np.ascontiguousarray(dir_, dtype=STREAM_ENTRY)
    meta = s.meta.to_numpy(s.nblocks) if s.nblocks else np.zeros(0, np.uint32)
    data = s.bytes_host() if s.nbytes else np.zeros(0, np.uint8)
    n = lib.utl_container_bytes(dir_.size, s.nblocks, s.nbytes)
    if n == 0:
        raise ValueError("container size overflows")
    out = np.empty(n, np.uint8)
    out_len = C.c_size_t()
    check(lib.utl_container_write(out.ctypes.data, n, dir_.ctypes.data, dir_.size, meta.ctypes.data, s.nblocks,
                                  data.ctypes.data, s.nbytes, C.byref(out_len)), "utl_container_write")
    return out


def parse_container(buf: np.ndarray):
    """Validates container bytes; returns (header record, directory, meta
    words, data bytes), the last three as views into buf."""
    lib = _lib.load()
    buf = np.asarray(buf, dtype=np.uint8)
    if buf.ctypes.data % 8:
        buf = buf.copy()  # sections are 8-byte aligned within the buffer
    hdr = np.zeros(1, CONTAINER_HEADER)
    pd, pm, px = C.c_void_p(), C.c_void_p(), C.c_void_p()
    check(lib.utl_container_parse(buf.ctypes.data, buf.size, hdr.ctypes.data, C.byref(pd), C.byref(pm),
                                  C.byref(px)), "utl_container_parse")
    h = hdr[0]
    do, mo, xo = int(h["dir_off"]), int(h["meta_off"]), int(h["data_off"])
    ns, nb, nx = int(h["nstreams"]), int(h["nblocks"]), int(h["nbytes"])
    dir_ = buf[do:do + ns * STREAM_ENTRY.itemsize].view(STREAM_ENTRY)
    meta = buf[mo:mo + 4 * nb].view(np.uint32)
    return h, dir_, meta, buf[xo:xo + nx]


def from_container(ctx: Context, buf: np.ndarray):
    """Uploads a container whose streams are all uint32 or all uint64 and
    builds its index on the device. Returns (DeviceStream, directory)."""
    h, dir_, meta, data = parse_container(buf)
    kinds = set(dir_["is_u64"].tolist())
    if len(kinds) > 1:
        raise ValueError("container mixes uint32 and uint64 streams; decode them by block range")
    is64 = bool(kinds.pop()) if kinds else False
    return from_host_blocks(ctx, data, meta & 0xFFFFFF, meta >> 24, is64), dir_.copy()


def save_container(path: str, s: DeviceStream, dir_: np.ndarray) -> int:
    buf = to_container(s, dir_)
    buf.tofile(path)
    return int(buf.size)


def load_container(ctx: Context, path: str):
    return from_container(ctx, np.fromfile(path, dtype=np.uint8))
