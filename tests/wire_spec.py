"""Independent pure-Python reader of the UTL-PFOR wire format.

Written from the format specification (utlpfor.ksy:7-67, README.md:14-17)
rather than from the decoder code, so that tests can check the oracle's
encoder layout against the spec without going through the oracle's own
decoder. Valid blocks only: no error semantics. Small inputs only.
"""
from __future__ import annotations

import struct


def parse_u32_block(buf: bytes, for64: bool = False, combine_ok: bool = False):
    """Parse one valid sub-block. Returns (values, header fields, length)."""
    (h,) = struct.unpack_from("<I", buf, 0)
    count = h & 0xFF
    bw = ((h >> 8) & 0x1F) * 4
    itype = (h >> 13) & 3
    forw = (h >> 15) & 3
    combine = bool(h & (1 << 19)) and combine_ok
    delta = bool(h & (1 << 22))
    zz = bool(h & (1 << 23))
    exc = (h >> 24) & 0xFF
    off = 4
    svb_len = 0
    if exc:
        (svb_len,) = struct.unpack_from("<H", buf, off)
        off += 2
    base = 0
    if forw:
        nb = 8 if (for64 and itype == 3 and not combine and forw == 3) else {1: 1, 2: 2, 3: 4}[forw]
        base = int.from_bytes(buf[off:off + nb], "little")
        off += nb
    block2_len = None
    if combine:
        (block2_len,) = struct.unpack_from("<H", buf, off)
        off += 2
    payload = buf[off:off + 16 * bw]
    off += 16 * bw
    vals = [0] * 128
    for i in range(128):
        lane, row = i % 16, i // 16
        bits = 0
        for b in range(bw):
            bitpos = row * bw + b
            word = bitpos // 32
            word_val = int.from_bytes(payload[word * 64 + lane * 4: word * 64 + lane * 4 + 4], "little")
            bits |= ((word_val >> (bitpos % 32)) & 1) << b
        vals[i] = bits
    if exc:
        if exc <= 16:
            positions = list(buf[off:off + exc])
            off += exc
        else:
            bitmap = int.from_bytes(buf[off:off + 16], "little")
            positions = [p for p in range(128) if bitmap >> p & 1]
            off += 16
        ctrl_n = (exc + 3) // 4
        svb = buf[off:off + svb_len]
        data = ctrl_n
        highs = []
        for i in range(exc):
            code = (svb[i // 4] >> (2 * (i % 4))) & 3
            highs.append(int.from_bytes(svb[data:data + code + 1], "little"))
            data += code + 1
        assert data == svb_len, "svb length must equal control + data bytes"
        for p, hb in zip(positions, highs):
            vals[p] |= (hb << bw) & 0xFFFFFFFF
        off += svb_len
    vals = vals[:count]
    if delta:
        for lane in range(16):
            acc = 0
            for row in range(8):
                i = row * 16 + lane
                if i >= count:
                    break
                d = vals[i]
                if zz:
                    d = (d >> 1) ^ (-(d & 1) & 0xFFFFFFFF)
                acc = (acc + d) & 0xFFFFFFFF if row else d & 0xFFFFFFFF
                vals[i] = acc
    if forw and not (for64 and itype == 3 and not combine and forw == 3):
        vals = [(v + base) & 0xFFFFFFFF for v in vals]
    fields = dict(count=count, bw=bw, itype=itype, forw=forw, delta=delta, zz=zz, exc=exc,
                  svb_len=svb_len, base=base, block2_len=block2_len, combine=combine)
    return vals, fields, off


def parse_u64_block(buf: bytes):
    vals, f, n = parse_u32_block(buf, for64=True, combine_ok=True)
    if f["combine"]:
        hi, f2, n2 = parse_u32_block(buf[n:n + f["block2_len"]], for64=True)
        return [(h << 32) | lo for lo, h in zip(vals, hi)], f, n + f["block2_len"]
    if f["itype"] == 3 and f["forw"] == 3:
        return [v + f["base"] for v in vals], f, n
    return vals, f, n
