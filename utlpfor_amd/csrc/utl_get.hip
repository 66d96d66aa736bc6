// utl_get.hip -- batched random access (GetUint32 / GetUint64) on gfx950.
//
// One thread serves one request. The per-thread path restates
// getUint32Scalar (get.go:81-190) and getUint64Scalar (uint64.go:372-436):
// direct lane-word extraction for non-delta blocks, a lane walk for delta
// blocks up to the full-unpack threshold of the reference's AVX-512 level
// (get.go:18-28: 4 rows for bw >= 20, else 8), and above it a full-block
// unpack with UnpackUint32's validation and overflow semantics evaluated for
// the requested position. StreamVByte single-value reads are clamped to the
// validated region (svb.go:40-95), so corrupt control bytes degrade to wrong
// values exactly as in the reference instead of reading out of bounds.
#include "utl_device.h"
#include "utl_format.h"
#include "utl_get.h"

namespace utl {

// svb.go:72 svbReadValue (reads clamped to len)
__device__ __forceinline__ uint32_t svb_read_value(const uint8_t* data, int64_t len, int64_t off, int bl) {
  if (off >= len) return 0;
  const int64_t end = off + bl < len ? off + bl : len;
  return ld_bytes(data + off, (int)(end - off));
}

// svb.go:40 svbDecodeOneInternal
__device__ uint32_t svb_decode_one(const uint8_t* enc, int64_t len, int count, int index) {
  if (count <= 0 || index < 0 || index >= count) return 0;
  int64_t data_off = svb_ctrl_count(count);
  const int target_cb = index >> 2, slot = index & 3;
  for (int cb = 0; cb < target_cb; ++cb) {
    if (cb >= len) return 0;
    data_off += svb_ctrl_block_size(ld_u8(enc + cb));
  }
  if (target_cb >= len) return 0;
  const uint32_t ctrl = ld_u8(enc + target_cb);
  for (int s = 0; s < slot; ++s) data_off += ((ctrl >> (s * 2)) & 3) + 1;
  return svb_read_value(enc, len, data_off, (int)((ctrl >> (slot * 2)) & 3) + 1);
}

// get.go:34 extractPackedValueUTL
__device__ __forceinline__ uint32_t extract_packed(int pos, const uint8_t* payload, int plen, int bw) {
  const int lane = pos % kLanes, row = pos / kLanes;
  const int bitpos = row * bw, word = bitpos / 32, bitoff = bitpos % 32;
  const int boff = word * kSuperWord + lane * 4;
  uint64_t acc = 0;
  if (boff + 4 <= plen) acc = ld_u32(payload + boff);
  if (bw > 32 - bitoff) {
    const int nb = boff + kSuperWord;
    if (nb + 4 <= plen) acc |= (uint64_t)ld_u32(payload + nb) << 32;
  }
  acc >>= bitoff;
  const uint64_t mask = bw >= 32 ? 0xFFFFFFFFull : ((1ull << bw) - 1);
  return (uint32_t)(acc & mask);
}

// get.go:363 bitmapRank128 (region of at least 16 bytes)
__device__ __forceinline__ int bitmap_rank128(const uint8_t* bmp, int64_t blen, int byte_idx, int bit_idx) {
  if (blen < 16) return 0;
  const uint64_t w0 = ld_u64(bmp), w1 = ld_u64(bmp + 8);
  if (byte_idx < 8) return __popcll(w0 & ((1ull << (8 * byte_idx + bit_idx)) - 1));
  return __popcll(w0) + __popcll(w1 & ((1ull << (8 * (byte_idx - 8) + bit_idx)) - 1));
}

// exceptions.go:293 findExceptionIndex over region = buf[excStart:]
__device__ int find_exception_index(const uint8_t* region, int64_t rlen, int exc, int pos) {
  if (exc <= kExcBitmap) {
    const int64_t n = exc < rlen ? exc : rlen;
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t b = ld_u8(region + i);
      if (b == (uint32_t)(uint8_t)pos) return (int)i;
      if (b > (uint32_t)(uint8_t)pos) return -1;
    }
    return -1;
  }
  const int64_t blen = rlen < 16 ? rlen : 16;
  if (blen == 0) return -1;
  const int byte_idx = pos / 8, bit_idx = pos % 8;
  if (byte_idx >= blen) return -1;
  if (!(ld_u8(region + byte_idx) & (1u << bit_idx))) return -1;
  return bitmap_rank128(region, blen, byte_idx, bit_idx);
}

// ---- the delta full-unpack fallback (get.go:339-349 / :198-218), 16 lanes a block ----
//
// UnpackUint32 semantics for one position: full validation of
// decodeUnpackHeader and the exception region, every value of the block
// decoded so that ErrOverflow reports the first wrap anywhere in it, then
// the requested value. A 16-lane group (one DPP row) decodes the block:
// lane L owns UTL lane L (the values at 16r + L), so each lane reads its 8
// payload words at once, patches its exceptions, runs its own per-lane
// delta, and the group takes the lowest wrap position (row-major order, the
// order of the reference's overflow report) with a 16-lane min.
// (buf, len, pos) are the same on every lane of the group; every lane
// returns the same status and value.

__device__ __forceinline__ uint32_t g16_add(uint32_t x) {
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) x += (uint32_t)__shfl_xor((int)x, d, 16);
  return x;
}
__device__ __forceinline__ uint32_t g16_min(uint32_t x) {
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) x = min(x, (uint32_t)__shfl_xor((int)x, d, 16));
  return x;
}
// exclusive prefix sum over the 16 lanes of a group
__device__ __forceinline__ uint32_t g16_excl_scan(uint32_t x, int L) {
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 16);
    if (L >= d) incl += y;
  }
  return incl - x;
}

template <bool ForU64>
__device__ int coop_unpack_at(const uint8_t* buf, int64_t len, int pos, int L, uint32_t* out) {
  *out = 0;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  const Hdr h = decode_header(ld_u32(buf));
  bool combine = h.combine;
  if (ForU64) {
    if (h.int_type != 3) return UTL_ERR_UNSUPPORTED_TYPE;
  } else {
    if (h.int_type != 2 && h.int_type != 1) return UTL_ERR_UNSUPPORTED_TYPE;
    combine = false;
  }
  if ((unsigned)(h.count - 1) >= (unsigned)kBlock || (unsigned)h.bw > 32u) {
    if (h.count == 0) return UTL_ERR_POSITION_OUT_OF_RANGE;  // unreachable: Get checked pos < count
    if (h.count > kBlock) return UTL_ERR_INVALID_BLOCK_LENGTH;
    return UTL_ERR_INVALID_BUFFER;
  }
  const bool has_exc = h.exc > 0, has_for = h.for_width > 0;
  const bool u64s = ForU64 && is_for64_single(h.int_type, h.for_width, combine);
  int poff = kHdr + (has_exc ? kSvbLen : 0);
  const int for_off = poff;
  if (has_for) poff += u64s ? kFor64 : for_base_bytes(h.for_width);
  if (combine) poff += kB2Len;
  const int pbytes = payload_bytes(h.bw);
  const int idx = has_exc ? exc_index_size(h.exc) : 0;
  const int64_t end = (int64_t)poff + pbytes + idx;
  if (len < end) return UTL_ERR_INVALID_BUFFER;
  int svb_len = 0;
  if (has_exc) {
    svb_len = (int)ld_u16(buf + kHdr);
    if (end + svb_len > len) return UTL_ERR_INVALID_BUFFER;
  }
  const uint32_t for_base = (has_for && !u64s) ? ld_bytes(buf + for_off, for_base_bytes(h.for_width)) : 0u;
  const int exc_start = poff + pbytes;
  const uint8_t* svb = buf + exc_start + idx;
  // lane L's control bytes 2L and 2L + 1 (their data-byte counts)
  uint32_t bs_a = 0, bs_b = 0, cb_excl = 0;
  if (has_exc) {
    if (h.exc > h.count) return UTL_ERR_INVALID_BUFFER;
    const int ctrl = svb_ctrl_count(h.exc);
    if (ctrl > svb_len) return UTL_ERR_INVALID_BUFFER;
    // data bytes the excCount codes announce: whole control bytes below
    // exc / 4, then the codes of the partial one
    const int full = h.exc >> 2, kk = h.exc & 3;
    if (2 * L < ctrl) bs_a = (uint32_t)svb_ctrl_block_size(ld_u8(svb + 2 * L));
    if (2 * L + 1 < ctrl) bs_b = (uint32_t)svb_ctrl_block_size(ld_u8(svb + 2 * L + 1));
    uint32_t mine = (2 * L < full ? bs_a : 0u) + (2 * L + 1 < full ? bs_b : 0u);
    int needed = ctrl + (int)g16_add(mine);
    if (kk > 0) {
      const uint32_t c = ld_u8(svb + full);
      for (int sl = 0; sl < kk; ++sl) needed += (int)((c >> (sl * 2)) & 3) + 1;
    }
    if (needed > svb_len) return UTL_ERR_INVALID_BUFFER;
    cb_excl = g16_excl_scan(bs_a + bs_b, L);
  }
  const uint8_t* payload = buf + poff;
  uint32_t v[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) v[r] = h.bw > 0 ? extract_packed(r * kLanes + L, payload, pbytes, h.bw) : 0u;
  if (has_exc) {
    if (h.exc <= kExcBitmap) {
      // exception k on lane k: its position, and its value from the
      // StreamVByte data at the sum of the lengths of the ones before
      const bool mine = L < h.exc;
      const uint32_t p = mine ? ld_u8(buf + exc_start + L) : 0u;
      const int tgt = !mine ? -1 : h.count == kBlock ? (int)(p & 0x7F) : ((int)p < h.count ? (int)p : -1);
      const uint32_t code = mine ? (ld_u8(svb + (L >> 2)) >> (2 * (L & 3))) & 3u : 0u;
      const uint32_t blen = mine ? code + 1 : 0u;
      const uint32_t off = (uint32_t)svb_ctrl_count(h.exc) + g16_excl_scan(blen, L);
      const uint32_t hv = mine ? go_shl32(svb_read_value(svb, svb_len, off, (int)blen), h.bw) : 0u;
      for (int k = 0; k < h.exc; ++k) {
        const int t = __shfl(tgt, k, 16);
        const uint32_t add = (uint32_t)__shfl((int)hv, k, 16);
        if (t >= 0 && (t & 15) == L) {
#pragma unroll
          for (int r = 0; r < kRows; ++r) v[r] |= r == (t >> 4) ? add : 0u;
        }
      }
    } else {
      uint64_t bw0 = ld_u64(buf + exc_start), bw1 = ld_u64(buf + exc_start + 8);
      if (h.count < kBlock) {
        if (h.count < 64) {
          bw1 = 0;
          bw0 &= (1ull << h.count) - 1;
        } else {
          bw1 &= (1ull << (h.count - 64)) - 1;
        }
      }
      if (__popcll(bw0) + __popcll(bw1) <= h.exc) {  // else applyBitmapExceptions patches nothing
        const int ctrl = svb_ctrl_count(h.exc);
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          const int q = r * kLanes + L;
          const uint64_t w = q < 64 ? bw0 : bw1;
          const bool set = (w >> (q & 63)) & 1;
          const int rank = (q < 64 ? 0 : __popcll(bw0)) + __popcll(w & ((1ull << (q & 63)) - 1));
          // data offset of value `rank`: the control bytes before its own
          // (prefix from the lane holding them), then the codes before it
          const int cb = rank >> 2;
          const uint32_t e = (uint32_t)__shfl((int)cb_excl, (cb >> 1) & 15, 16);
          const uint32_t a = (uint32_t)__shfl((int)bs_a, (cb >> 1) & 15, 16);
          if (set) {
            const uint32_t c = ld_u8(svb + cb);
            uint32_t off = (uint32_t)ctrl + e + ((cb & 1) ? a : 0u);
            for (int sl = 0; sl < (rank & 3); ++sl) off += ((c >> (2 * sl)) & 3) + 1;
            const int bl = (int)((c >> (2 * (rank & 3))) & 3) + 1;
            v[r] |= go_shl32(svb_read_value(svb, svb_len, off, bl), h.bw);
          }
        }
      }
    }
  }
  uint32_t wrap = 0xFFFFFFFFu;  // this lane's first wrap position
  if (h.delta) {
    uint32_t prev = 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int q = r * kLanes + L;
      if (q < h.count) {
        if (h.zigzag) {
          v[r] = r == 0 ? zz_dec(v[r]) : prev + zz_dec(v[r]);
        } else if (r > 0) {
          const uint32_t nx = prev + v[r];
          if (nx < prev && wrap == 0xFFFFFFFFu) wrap = (uint32_t)q;
          v[r] = nx;
        }
      }
      prev = v[r];
    }
  }
  const uint32_t ovf = g16_min(wrap);
  uint32_t sel = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) sel = r == (pos >> 4) ? v[r] : sel;
  const uint32_t result = (uint32_t)__shfl((int)sel, pos & 15, 16);
  if (h.delta && !h.zigzag && ovf != 0xFFFFFFFFu) return UTL_ERR_OVERFLOW | ((int)ovf << 8);
  *out = result + ((has_for && !u64s) ? for_base : 0u);
  return UTL_OK;
}

// A Get handed to the 16-lane full unpack (coop_unpack_at) of its block.
constexpr int kGetDefer = -1;

// get.go:81 getUint32Scalar. Returns kGetDefer, with *dlen the block's
// validated extent, where the reference takes the full-unpack fallback.
template <bool ForU64>
__device__ int get_u32(int64_t pos, const uint8_t* src, int64_t len, uint32_t* out, int64_t* dlen) {
  *out = 0;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  const Hdr h = decode_header(ld_u32(src));
  const bool has_for = h.for_width > 0;
  const bool has_exc = h.exc > 0;
  bool combine = h.combine;
  if (ForU64) {
    if (h.int_type != 3) return UTL_ERR_UNSUPPORTED_TYPE;
  } else {
    if (h.int_type != 2 && h.int_type != 1) return UTL_ERR_UNSUPPORTED_TYPE;
    combine = false;
  }
  if (pos < 0 || pos >= h.count) return UTL_ERR_POSITION_OUT_OF_RANGE;
  if (h.bw > 32) return UTL_ERR_INVALID_BUFFER;
  const bool u64s = ForU64 && is_for64_single(h.int_type, h.for_width, combine);
  int poff = kHdr + (has_exc ? kSvbLen : 0);
  int forb = 0;
  if (has_for) {
    if (u64s) {
      poff += kFor64;
    } else {
      forb = for_base_bytes(h.for_width);
      poff += forb;
    }
  }
  if (combine) poff += kB2Len;
  const int pbytes = payload_bytes(h.bw);
  int64_t end = (int64_t)poff + pbytes;
  if (len < end) return UTL_ERR_INVALID_BUFFER;
  if (has_exc) {
    if (h.exc > h.count) return UTL_ERR_INVALID_BUFFER;
    const int svb_len = (int)ld_u16(src + kHdr);
    const int isz = exc_index_size(h.exc);
    if (svb_len < svb_ctrl_count(h.exc) || end + isz + svb_len > len) return UTL_ERR_INVALID_BUFFER;
    end += isz + svb_len;
  }
  len = end;
  const uint32_t for_base = forb ? ld_bytes(src + poff - forb, forb) : 0u;
  const uint8_t* payload = src + poff;
  const int exc_start = poff + pbytes;
  const int p = (int)pos;
  if (!h.delta) {
    // get.go:228 getValueDirect
    uint32_t value = h.bw > 0 ? extract_packed(p, payload, pbytes, h.bw) : 0u;
    if (has_exc) {
      const uint8_t* region = src + exc_start;
      const int64_t rlen = len - exc_start;
      const int ei = find_exception_index(region, rlen, h.exc, p);
      if (ei >= 0) {
        const int isz = exc_index_size(h.exc);
        value |= go_shl32(svb_decode_one(region + isz, rlen - isz, h.exc, ei), h.bw);
      }
    }
    *out = value + for_base;
    return UTL_OK;
  }
  const int row = p / kLanes, lane = p % kLanes;
  const int threshold = h.bw >= 20 ? 4 : kRows;
  if (row > threshold) {
    *dlen = len;
    return kGetDefer;
  }
  // The reference indexes its 8-entry lane array at row 8 here (count > 128)
  // and panics; this implementation reports ErrInvalidBuffer instead.
  if (row >= kRows) return UTL_ERR_INVALID_BUFFER;
  // get.go:250 getValueWithDelta
  uint32_t lv[kRows];
#pragma unroll
  for (int v = 0; v < kRows; ++v) lv[v] = 0;
  for (int v = 0; v <= row; ++v) {
    const int seq = lane + v * kLanes;
    if (seq >= h.count) break;
    if (h.bw > 0) lv[v] = extract_packed(seq, payload, pbytes, h.bw);
  }
  if (has_exc) {
    // get.go:297 applyLaneExceptions
    const uint8_t* region = src + exc_start;
    const int64_t rlen = len - exc_start;
    const int isz = exc_index_size(h.exc);
    const uint8_t* svb = region + isz;
    const int64_t svb_len = rlen - isz;
    if (h.exc <= kExcBitmap) {
      for (int i = 0; i < h.exc; ++i) {
        const int q = (int)ld_u8(region + i);
        if (q % kLanes != lane) continue;
        const int v = q / kLanes;
        if (v > row) break;
        lv[v] |= go_shl32(svb_decode_one(svb, svb_len, h.exc, i), h.bw);
      }
    } else {
      for (int v = 0; v <= row; ++v) {
        const int seq = lane + v * kLanes;
        if (seq >= h.count) break;
        const int byte_idx = seq / 8, bit_idx = seq % 8;
        if (!(ld_u8(region + byte_idx) & (1u << bit_idx))) continue;
        const int rank = bitmap_rank128(region, 16, byte_idx, bit_idx);
        lv[v] |= go_shl32(svb_decode_one(svb, svb_len, h.exc, rank), h.bw);
      }
    }
  }
  if (h.zigzag) lv[0] = zz_dec(lv[0]);
  for (int v = 1; v <= row; ++v) {
    const int seq = lane + v * kLanes;
    if (seq >= h.count) break;
    lv[v] = lv[v - 1] + (h.zigzag ? zz_dec(lv[v]) : lv[v]);
  }
  *out = lv[row] + for_base;
  return UTL_OK;
}

// One request's Get as a state machine, so that the full-unpack fallback
// of any sub-block Get can run on a 16-lane group of the wavefront and the
// request then continue where it stopped. uint32 requests: one sub-Get.
// uint64 requests follow getUint64Scalar (uint64.go:372-436): the lower
// sub-block's Get; then, without combine, the FOR64 base; with combine,
// block1Len / block2Len checks and the upper sub-block's Get.
struct GetJob {
  const uint8_t* src;   // the request's block
  int64_t len;
  int64_t pos;
  int stage;            // kJobDone or one of the waits below
  int status;
  uint64_t value;       // the result; while the upper sub-block waits, the lower half
  // the pending sub-block Get (stage != kJobDone)
  const uint8_t* sub;
  int64_t sublen;
};
constexpr int kJobDone = 0, kJobU32 = 1, kJobNcLower = 2, kJobLower = 3, kJobUpper = 4;

__device__ __forceinline__ void job_done(GetJob& j, int st, uint64_t v) {
  j.stage = kJobDone;
  j.status = st;
  j.value = st == UTL_OK ? v : 0ull;
}

// uint64: after the non-combine sub-block Get (FOR64 base, uint64.go:389-399)
__device__ __forceinline__ void job_nc_lower(GetJob& j, int err, uint32_t lower) {
  if (err) return job_done(j, err, 0);
  const Hdr h = decode_header(ld_u32(j.src));
  uint64_t base = 0;
  if (is_for64_single(h.int_type, h.for_width, h.combine)) {
    const int boff = kHdr + (h.exc > 0 ? kSvbLen : 0);
    if (j.len < boff + kFor64) return job_done(j, UTL_ERR_INVALID_BUFFER, 0);
    base = ld_u64(j.src + boff);
  }
  job_done(j, UTL_OK, (uint64_t)lower + base);
}

__device__ __forceinline__ void job_upper(GetJob& j, int err, uint32_t upper) {
  if (err) return job_done(j, err, 0);
  job_done(j, UTL_OK, ((uint64_t)upper << 32) | (uint32_t)j.value);
}

// uint64 combine: after the lower sub-block Get (uint64.go:408-436)
__device__ __forceinline__ void job_lower(GetJob& j, int err, uint32_t lower) {
  if (err) return job_done(j, err, 0);
  const Hdr h = decode_header(ld_u32(j.src));
  const bool has_exc = h.exc > 0;
  const int forb = for_base_bytes(h.for_width);
  const int meta = payload_offset(forb, has_exc, true);
  if (j.len < meta) return job_done(j, UTL_ERR_INVALID_BUFFER, 0);
  const int b2len = (int)ld_u16(j.src + kHdr + (has_exc ? kSvbLen : 0) + forb);
  int64_t b1len = meta + (h.bw << 4);
  if (has_exc) {
    if (j.len < kHdr + kSvbLen) return job_done(j, UTL_ERR_INVALID_BUFFER, 0);
    b1len += (h.exc < kExcBitmap ? h.exc : kExcBitmap) + (int64_t)ld_u16(j.src + kHdr);
  }
  if (b1len + b2len > j.len) return job_done(j, UTL_ERR_INVALID_BUFFER, 0);
  j.value = lower;
  uint32_t upper;
  int64_t dlen;
  const int st = get_u32<true>(j.pos, j.src + b1len, b2len, &upper, &dlen);
  if (st == kGetDefer) {
    j.stage = kJobUpper;
    j.sub = j.src + b1len;
    j.sublen = dlen;
    return;
  }
  job_upper(j, st, upper);
}

// Starts request j (its block at src, len bytes, position pos).
template <bool U64>
__device__ __forceinline__ void job_start(GetJob& j, const uint8_t* src, int64_t len, int64_t pos) {
  j.src = src;
  j.len = len;
  j.pos = pos;
  uint32_t v;
  int64_t dlen;
  if (!U64) {
    const int st = get_u32<false>(pos, src, len, &v, &dlen);
    if (st == kGetDefer) {
      j.stage = kJobU32;
      j.sub = src;
      j.sublen = dlen;
      return;
    }
    return job_done(j, st, v);
  }
  if (len < kHdr) return job_done(j, UTL_ERR_INVALID_BUFFER, 0);
  const Hdr h = decode_header(ld_u32(src));
  if (h.int_type != 3) return job_done(j, UTL_ERR_UNSUPPORTED_TYPE, 0);
  const int st = get_u32<true>(pos, src, len, &v, &dlen);
  if (st == kGetDefer) {
    j.stage = h.combine ? kJobLower : kJobNcLower;
    j.sub = src;
    j.sublen = dlen;
    return;
  }
  if (h.combine) job_lower(j, st, v);
  else job_nc_lower(j, st, v);
}

// Continues request j with the result of its pending sub-block Get.
__device__ __forceinline__ void job_resume(GetJob& j, int st, uint32_t v) {
  switch (j.stage) {
    case kJobU32: job_done(j, st, v); break;
    case kJobNcLower: job_nc_lower(j, st, v); break;
    case kJobLower: job_lower(j, st, v); break;
    case kJobUpper: job_upper(j, st, v); break;
    default: break;
  }
}

// A position-directory entry (k_pos_dir): the block b holding position
// 128c and what a Get needs of blocks b and b + 1 without reading the offset
// arrays. lens: boff[b + 1] - boff[b] | (boff[b + 2] - boff[b + 1]) << 16;
// cnts: ooff[b + 1] - ooff[b] | (ooff[b + 2] - ooff[b + 1]) << 8 |
// kDirUse0 when block b's fields fit | kDirUse1 when block b + 1's do too.
struct alignas(16) PosDirEntry {
  uint64_t boff0, ooff0, b;
  uint32_t lens, cnts;
};
static_assert(sizeof(PosDirEntry) == 4 * sizeof(uint64_t), "four words per directory entry");
constexpr uint32_t kDirUse0 = 1u << 16, kDirUse1 = 1u << 17;

// Requests: either (req_block[i], req_pos[i]) pairs, or stream value
// positions req_pos[i], each mapped to the last block b with ooff[b] <= pos.
// With a position directory (dir != nullptr, ndir entries of PosDirEntry)
// a position in the entry's block or the next one takes that block's
// offsets from the entry (with a nondecreasing ooff, ooff[b] <= pos <
// ooff[b + 1] makes b the last block whose entry is <= pos); otherwise the
// search gallops from the entry's block. Without one it gallops from
// pos / 128.
// Two launches over the same requests. The first (FULL false) runs each
// request's Get on one thread and marks the requests that reach the
// reference's full-unpack fallback with status kGetDeferStatus. The second
// (FULL true) runs only the marked ones (or, with `all`, every request, in
// one launch for small batches): their Get again up to that point, then
// the fallback in rounds, four requests at a time, each on a 16-lane group
// of the wavefront (every lane of a wave with such requests takes part, so
// none leaves before the rounds end). Keeping the fallback's registers out
// of the first launch keeps its occupancy for the gathers.
constexpr int32_t kGetDeferStatus = -1;

template <bool U64, bool FULL>
__global__ __launch_bounds__(256) void k_get(const uint8_t* __restrict__ src, const uint64_t* __restrict__ boff,
                                             const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                             const PosDirEntry* __restrict__ dir, uint64_t ndir,
                                             const uint64_t* __restrict__ req_block, const int64_t* __restrict__ req_pos,
                                             uint64_t nreq, void* __restrict__ out, int32_t* __restrict__ status,
                                             bool all) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < nreq && (!FULL || all || status[i] == kGetDeferStatus);
  if (FULL && !__ballot(live)) return;  // (the whole wave)
  GetJob job;
  job.stage = kJobDone;
  job.status = UTL_OK;
  job.value = 0;
  job.sub = src;
  job.sublen = 0;
  job.pos = 0;
  if (live) {
    int64_t pos = req_pos[i];
    uint64_t b;
    int st;
    // the block's bytes and length when the directory entry gave them
    const uint8_t* blk = nullptr;
    int64_t blen = 0;
    if (req_block) {
      b = req_block[i];
      st = b < nblocks ? UTL_OK : UTL_ERR_ARGUMENT;
    } else {
      st = (pos >= 0 && (uint64_t)pos < ooff[nblocks]) ? UTL_OK : UTL_ERR_POSITION_OUT_OF_RANGE;
      // Largest b with ooff[b] <= pos: a bracket [lo, hi) with ooff[lo] <=
      // pos < ooff[hi] (hi == nblocks: past the end; lo == 0 also when no
      // entry is <= pos), then bisection. Any nondecreasing out_off gives
      // the same b as a bisection of the whole array.
      uint64_t lo = 0, hi = nblocks;
      if (dir) {
        // entry c's block is the largest b with ooff[b] <= 128c (or 0), and
        // 128c <= pos: the entry's block or the next, else gallop up from it
        const uint64_t c = st == UTL_OK ? (uint64_t)pos >> 7 : 0;
        if (c < ndir) {
          const PosDirEntry e = dir[c];
          const uint64_t up = (uint64_t)pos, c0 = e.cnts & 0xFFu, c1 = (e.cnts >> 8) & 0xFFu;
          if (st == UTL_OK && up >= e.ooff0 && (e.cnts & kDirUse0) && up < e.ooff0 + c0) {
            b = e.b;
            blk = src + e.boff0;
            blen = (int64_t)(e.lens & 0xFFFFu);
            pos = (int64_t)(up - e.ooff0);
          } else if (st == UTL_OK && up >= e.ooff0 && (e.cnts & kDirUse1) && up < e.ooff0 + c0 + c1) {
            b = e.b + 1;
            blk = src + e.boff0 + (e.lens & 0xFFFFu);
            blen = (int64_t)(e.lens >> 16);
            pos = (int64_t)(up - e.ooff0 - c0);
          }
          lo = min(e.b, nblocks - 1);
        }
        if (!blk) {
          for (uint64_t step = 1; lo + step < nblocks; step <<= 1) {
            if (ooff[lo + step] > (uint64_t)pos) {
              hi = lo + step;
              break;
            }
            lo += step;
          }
        }
      } else {
        // Blocks of an index built from meta words hold at most 128 values,
        // so b >= pos / 128, and b == pos / 128 when the blocks before are
        // full: gallop from there.
        const uint64_t g = st == UTL_OK ? min((uint64_t)pos >> 7, nblocks - 1) : 0;
        if (ooff[g] <= (uint64_t)pos) {
          lo = g;
          for (uint64_t step = 1; lo + step < nblocks; step <<= 1) {
            if (ooff[lo + step] > (uint64_t)pos) {
              hi = lo + step;
              break;
            }
            lo += step;
          }
        } else {
          hi = g;
          for (uint64_t step = 1; step <= hi; step <<= 1) {
            if (ooff[hi - step] <= (uint64_t)pos) {
              lo = hi - step;
              break;
            }
            hi -= step;
          }
        }
      }
      if (!blk) {
        while (hi - lo > 1) {
          const uint64_t mid = (lo + hi) >> 1;
          if (ooff[mid] <= (uint64_t)pos) lo = mid;
          else hi = mid;
        }
        b = lo;
        pos -= (int64_t)ooff[b];
      }
    }
    if (st == UTL_OK) {
      if (!blk) {
        blk = src + boff[b];
        blen = (int64_t)(boff[b + 1] - boff[b]);
      }
      job_start<U64>(job, blk, blen, pos);
    } else {
      job_done(job, st, 0);
    }
  }
  if (!FULL) {
    if (live) {
      const bool wait = job.stage != kJobDone;
      if (U64) ((uint64_t*)out)[i] = wait ? 0ull : job.value;
      else ((uint32_t*)out)[i] = wait ? 0u : (uint32_t)job.value;
      status[i] = wait ? kGetDeferStatus : job.status;
    }
    return;
  }
  // full-unpack rounds: group g (lanes 16g .. 16g + 15) serves the g-th
  // lowest waiting lane; served lanes take the group's result and go on
  const int lane = lane_id(), g = lane >> 4, L = lane & 15;
  for (;;) {
    const uint64_t wait = __ballot(job.stage != kJobDone);
    if (!wait) break;
    uint64_t m = wait;
    int srv = -1;  // the lane this lane's group serves
    for (int k = 0; k < 4 && m; ++k) {
      const int l = __builtin_ctzll(m);
      if (k == g) srv = l;
      m &= m - 1;
    }
    const int from = srv < 0 ? lane : srv;
    const uint64_t sp = ((uint64_t)(uint32_t)__shfl((int)(reinterpret_cast<uintptr_t>(job.sub) >> 32), from, 64) << 32) |
                        (uint32_t)__shfl((int)(uint32_t)reinterpret_cast<uintptr_t>(job.sub), from, 64);
    const int64_t sl = ((int64_t)__shfl((int)((uint64_t)job.sublen >> 32), from, 64) << 32) |
                       (int64_t)(uint32_t)__shfl((int)(uint32_t)job.sublen, from, 64);
    const int sposv = __shfl((int)job.pos, from, 64);
    int cst = UTL_OK;
    uint32_t cval = 0;
    if (srv >= 0) cst = coop_unpack_at<U64>(reinterpret_cast<const uint8_t*>(sp), sl, sposv, L, &cval);
    // a waiting lane ranked below 4 among the waiting lanes was served by
    // group `rank`, whose lane 16 * rank holds the result
    const int rank = __popcll(wait & ((1ull << lane) - 1));
    const int gst = __shfl(cst, (rank & 3) << 4, 64);
    const uint32_t gval = (uint32_t)__shfl((int)cval, (rank & 3) << 4, 64);
    if (job.stage != kJobDone && rank < 4) job_resume(job, gst, gval);
  }
  if (live) {
    if (U64) ((uint64_t*)out)[i] = job.value;
    else ((uint32_t*)out)[i] = (uint32_t)job.value;
    status[i] = job.status;
  }
}

// Position directory: entry c describes the largest block b with ooff[b]
// <= 128c (0 when there is none), for c < ndir. Thread b writes the entries
// c with 128c in [ooff[b], ooff[b + 1]), widened to [0, ooff[1]) for block 0
// and to [ooff[b], 128 ndir) for the last block; with a nondecreasing ooff
// these ranges tile [0, 128 ndir), so every entry is written exactly once,
// and the block of each range is the largest whose offset is <= any
// position in it (block 0 also where no offset is). A block's fields are
// marked usable when its byte length and value count fit them (differences
// of the offset arrays, as unsigned numbers).
__global__ __launch_bounds__(256) void k_pos_dir(const uint64_t* __restrict__ boff, const uint64_t* __restrict__ ooff,
                                                 uint64_t nblocks, PosDirEntry* __restrict__ dir, uint64_t ndir) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint64_t lo = b == 0 ? 0 : ooff[b], hi = b + 1 == nblocks ? (ndir << 7) : ooff[b + 1];
  uint64_t c = (lo + 127) >> 7;
  if (c >= ndir || (c << 7) >= hi) return;
  PosDirEntry e;
  e.boff0 = boff[b];
  e.ooff0 = ooff[b];
  e.b = b;
  const uint64_t l0 = boff[b + 1] - e.boff0, n0 = ooff[b + 1] - e.ooff0;
  const bool use0 = l0 <= 0xFFFFu && n0 <= 0xFFu;
  uint64_t l1 = 0, n1 = 0;
  if (b + 1 < nblocks) {
    l1 = boff[b + 2] - boff[b + 1];
    n1 = ooff[b + 2] - ooff[b + 1];
  }
  const bool use1 = use0 && b + 1 < nblocks && l1 <= 0xFFFFu && n1 <= 0xFFu;
  e.lens = (use0 ? (uint32_t)l0 : 0u) | (use1 ? (uint32_t)l1 << 16 : 0u);
  e.cnts = (use0 ? (uint32_t)n0 | kDirUse0 : 0u) | (use1 ? ((uint32_t)n1 << 8) | kDirUse1 : 0u);
  for (; c < ndir && (c << 7) < hi; ++c) dir[c] = e;
}

}  // namespace utl

using namespace utl;

hipError_t utl_launch_get(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                          const uint64_t* dir_words, uint64_t ndir, const uint64_t* req_block, const int64_t* req_pos,
                          uint64_t nreq, void* out, int32_t* status, hipStream_t stream) {
  if (nreq == 0) return hipSuccess;
  const PosDirEntry* dir = reinterpret_cast<const PosDirEntry*>(dir_words);
  const dim3 grid((unsigned)((nreq + 255) / 256)), block(256);
  // one workgroup's worth: the fallback kernel alone takes every request
  const bool one = nreq <= 256;
  if (u64) {
    if (!one)
      k_get<true, false><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, dir, ndir, req_block, req_pos, nreq,
                                                     out, status, false);
    k_get<true, true><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, dir, ndir, req_block, req_pos, nreq, out,
                                                  status, one);
  } else {
    if (!one)
      k_get<false, false><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, dir, ndir, req_block, req_pos, nreq,
                                                      out, status, false);
    k_get<false, true><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, dir, ndir, req_block, req_pos, nreq, out,
                                                   status, one);
  }
  return hipGetLastError();
}

hipError_t utl_launch_pos_dir(const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks, uint64_t* dir_words,
                              uint64_t ndir, hipStream_t stream) {
  if (nblocks == 0 || ndir == 0) return hipSuccess;
  k_pos_dir<<<dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, stream>>>(
      boff, ooff, nblocks, reinterpret_cast<PosDirEntry*>(dir_words), ndir);
  return hipGetLastError();
}
