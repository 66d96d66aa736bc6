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

// UnpackUint32 semantics for one position (the delta full-unpack fallback of
// get.go:339-349 / :198-218): full validation of decodeUnpackHeader and the
// exception region, every lane decoded so that ErrOverflow is reported for
// the first wrap anywhere in the block, then the requested value.
template <bool ForU64>
__device__ int unpack_value_at(const uint8_t* buf, int64_t len, int pos, uint32_t* out) {
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
  int64_t end = (int64_t)poff + pbytes + idx;
  if (len < end) return UTL_ERR_INVALID_BUFFER;
  int svb_len = 0;
  if (has_exc) {
    svb_len = (int)ld_u16(buf + kHdr);
    if (end + svb_len > len) return UTL_ERR_INVALID_BUFFER;
  }
  const uint32_t for_base = (has_for && !u64s) ? ld_bytes(buf + for_off, for_base_bytes(h.for_width)) : 0u;
  const int exc_start = poff + pbytes;
  const int svb_start = exc_start + idx;
  uint64_t bw0 = 0, bw1 = 0;
  if (has_exc) {
    if (h.exc > h.count) return UTL_ERR_INVALID_BUFFER;
    const int ctrl = svb_ctrl_count(h.exc);
    if (ctrl > svb_len) return UTL_ERR_INVALID_BUFFER;
    int needed = ctrl;
    for (int g = 0; g < h.exc / 4; ++g) needed += svb_ctrl_block_size(ld_u8(buf + svb_start + g));
    const int kk = h.exc % 4;
    if (kk > 0) {
      const uint32_t c = ld_u8(buf + svb_start + h.exc / 4);
      for (int s = 0; s < kk; ++s) needed += ((c >> (s * 2)) & 3) + 1;
    }
    if (needed > svb_len) return UTL_ERR_INVALID_BUFFER;
    if (h.exc > kExcBitmap) {
      bw0 = ld_u64(buf + exc_start);
      bw1 = ld_u64(buf + exc_start + 8);
      if (h.count < kBlock) {
        if (h.count < 64) {
          bw1 = 0;
          bw0 &= (1ull << h.count) - 1;
        } else {
          bw1 &= (1ull << (h.count - 64)) - 1;
        }
      }
      if (__popcll(bw0) + __popcll(bw1) > h.exc) bw0 = bw1 = 0;  // applyBitmapExceptions skips
    }
  }
  const uint8_t* payload = buf + poff;
  const uint8_t* svb = buf + svb_start;
  const int64_t svb_region = svb_len;
  int ovf = 0;
  uint32_t result = 0;
  for (int lane = 0; lane < kLanes; ++lane) {
    uint32_t prev = 0;
    for (int r = 0; r < kRows; ++r) {
      const int q = r * kLanes + lane;
      uint32_t v = h.bw > 0 ? extract_packed(q, payload, pbytes, h.bw) : 0u;
      if (has_exc) {
        if (h.exc <= kExcBitmap) {
          for (int i = 0; i < h.exc; ++i) {
            const uint32_t p = ld_u8(buf + exc_start + i);
            const int tgt = h.count == kBlock ? (int)(p & 0x7F) : ((int)p < h.count ? (int)p : -1);
            if (tgt == q) v |= go_shl32(svb_decode_one(svb, svb_region, h.exc, i), h.bw);
          }
        } else {
          const uint64_t w = q < 64 ? bw0 : bw1;
          if ((w >> (q & 63)) & 1) {
            const int rank = (q < 64 ? 0 : __popcll(bw0)) + __popcll(w & ((1ull << (q & 63)) - 1));
            v |= go_shl32(svb_decode_one(svb, svb_region, h.exc, rank), h.bw);
          }
        }
      }
      if (h.delta && q < h.count) {
        if (h.zigzag) {
          v = r == 0 ? zz_dec(v) : prev + zz_dec(v);
        } else if (r > 0) {
          const uint32_t nx = prev + v;
          if (nx < prev && (ovf == 0 || q < ovf)) ovf = q;
          v = nx;
        }
      }
      prev = v;
      if (q == pos) result = v;
    }
  }
  if (h.delta && !h.zigzag && ovf > 0) return UTL_ERR_OVERFLOW | (ovf << 8);
  if (has_for && !u64s) result += for_base;
  *out = result;
  return UTL_OK;
}

// get.go:81 getUint32Scalar
template <bool ForU64>
__device__ int get_u32(int64_t pos, const uint8_t* src, int64_t len, uint32_t* out) {
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
  if (row > threshold) return unpack_value_at<ForU64>(src, len, p, out);
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

// uint64.go:372 getUint64Scalar
__device__ int get_u64(int64_t pos, const uint8_t* src, int64_t len, uint64_t* out) {
  *out = 0;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  const Hdr h = decode_header(ld_u32(src));
  if (h.int_type != 3) return UTL_ERR_UNSUPPORTED_TYPE;
  const bool has_exc = h.exc > 0;
  uint32_t lower, upper;
  if (!h.combine) {
    const int err = get_u32<true>(pos, src, len, &lower);
    if (err) return err;
    uint64_t base = 0;
    if (is_for64_single(h.int_type, h.for_width, h.combine)) {
      const int boff = kHdr + (has_exc ? kSvbLen : 0);
      if (len < boff + kFor64) return UTL_ERR_INVALID_BUFFER;
      base = ld_u64(src + boff);
    }
    *out = (uint64_t)lower + base;
    return UTL_OK;
  }
  int err = get_u32<true>(pos, src, len, &lower);
  if (err) return err;
  const int forb = for_base_bytes(h.for_width);
  const int meta = payload_offset(forb, has_exc, true);
  if (len < meta) return UTL_ERR_INVALID_BUFFER;
  const int b2len = (int)ld_u16(src + kHdr + (has_exc ? kSvbLen : 0) + forb);
  int64_t b1len = meta + (h.bw << 4);
  if (has_exc) {
    if (len < kHdr + kSvbLen) return UTL_ERR_INVALID_BUFFER;
    b1len += (h.exc < kExcBitmap ? h.exc : kExcBitmap) + (int64_t)ld_u16(src + kHdr);
  }
  if (b1len + b2len > len) return UTL_ERR_INVALID_BUFFER;
  err = get_u32<true>(pos, src + b1len, b2len, &upper);
  if (err) return err;
  *out = ((uint64_t)upper << 32) | lower;
  return UTL_OK;
}

// Requests: either (req_block[i], req_pos[i]) pairs, or stream value
// positions req_pos[i], each mapped to the last block b with ooff[b] <= pos.
// With a position directory (dir != nullptr, see k_pos_dir) the search
// starts at dir[pos / 128], the block holding position pos & ~127, so it
// usually reads one out_off line; without one it gallops from pos / 128.
template <bool U64>
__global__ __launch_bounds__(256) void k_get(const uint8_t* __restrict__ src, const uint64_t* __restrict__ boff,
                                             const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                             const uint64_t* __restrict__ dir, uint64_t ndir,
                                             const uint64_t* __restrict__ req_block, const int64_t* __restrict__ req_pos,
                                             uint64_t nreq, void* __restrict__ out, int32_t* __restrict__ status) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nreq) return;
  int64_t pos = req_pos[i];
  uint64_t b;
  int st;
  if (req_block) {
    b = req_block[i];
    st = b < nblocks ? UTL_OK : UTL_ERR_ARGUMENT;
  } else {
    st = (pos >= 0 && (uint64_t)pos < ooff[nblocks]) ? UTL_OK : UTL_ERR_POSITION_OUT_OF_RANGE;
    // Largest b with ooff[b] <= pos: a bracket [lo, hi) with ooff[lo] <=
    // pos < ooff[hi] (hi == nblocks: past the end; lo == 0 also when no
    // entry is <= pos), then bisection. Any nondecreasing out_off gives the
    // same b as a bisection of the whole array.
    uint64_t lo = 0, hi = nblocks;
    if (dir) {
      // dir[c] is the largest b with ooff[b] <= 128c (or 0), and 128c <= pos:
      // gallop up from it
      const uint64_t c = st == UTL_OK ? (uint64_t)pos >> 7 : 0;
      lo = c < ndir ? min(dir[c], nblocks - 1) : 0;
      for (uint64_t step = 1; lo + step < nblocks; step <<= 1) {
        if (ooff[lo + step] > (uint64_t)pos) {
          hi = lo + step;
          break;
        }
        lo += step;
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
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (ooff[mid] <= (uint64_t)pos) lo = mid;
      else hi = mid;
    }
    b = lo;
    pos -= (int64_t)ooff[b];
  }
  if (st == UTL_OK) {
    const uint8_t* p = src + boff[b];
    const int64_t len = (int64_t)(boff[b + 1] - boff[b]);
    if (U64) {
      uint64_t v;
      st = get_u64(pos, p, len, &v);
      ((uint64_t*)out)[i] = st == UTL_OK ? v : 0ull;
    } else {
      uint32_t v;
      st = get_u32<false>(pos, p, len, &v);
      ((uint32_t*)out)[i] = st == UTL_OK ? v : 0u;
    }
  } else {
    if (U64) ((uint64_t*)out)[i] = 0;
    else ((uint32_t*)out)[i] = 0;
  }
  status[i] = st;
}

// Position directory: dir[c] = the largest block b with ooff[b] <= 128c
// (0 when there is none), for c < ndir. Thread b writes the entries c with
// 128c in [ooff[b], ooff[b + 1]), widened to [0, ooff[1]) for block 0 and
// to [ooff[b], 128 ndir) for the last block; with a nondecreasing ooff these
// ranges tile [0, 128 ndir), so every entry is written exactly once, and
// the block of each range is the largest whose offset is <= any position in
// it (block 0 also where no offset is).
__global__ __launch_bounds__(256) void k_pos_dir(const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                                 uint64_t* __restrict__ dir, uint64_t ndir) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const uint64_t lo = b == 0 ? 0 : ooff[b], hi = b + 1 == nblocks ? (ndir << 7) : ooff[b + 1];
  for (uint64_t c = (lo + 127) >> 7; c < ndir && (c << 7) < hi; ++c) dir[c] = b;
}

}  // namespace utl

using namespace utl;

hipError_t utl_launch_get(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                          const uint64_t* dir, uint64_t ndir, const uint64_t* req_block, const int64_t* req_pos,
                          uint64_t nreq, void* out, int32_t* status, hipStream_t stream) {
  if (nreq == 0) return hipSuccess;
  const dim3 grid((unsigned)((nreq + 255) / 256)), block(256);
  if (u64)
    k_get<true><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, dir, ndir, req_block, req_pos, nreq, out, status);
  else
    k_get<false><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, dir, ndir, req_block, req_pos, nreq, out, status);
  return hipGetLastError();
}

hipError_t utl_launch_pos_dir(const uint64_t* ooff, uint64_t nblocks, uint64_t* dir, uint64_t ndir,
                              hipStream_t stream) {
  if (nblocks == 0 || ndir == 0) return hipSuccess;
  k_pos_dir<<<dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, stream>>>(ooff, nblocks, dir, ndir);
  return hipGetLastError();
}
