// utl_decode.hip -- gfx950 decode kernels for UTL-PFOR blocks.
//
// One wavefront decodes one 128-value block. Thread t (0..63) owns the two
// values at flat positions 2t and 2t+1, i.e. row r = t >> 3 and the lane
// pair (2k, 2k+1) with k = t & 7. Lane L's 32-bit words sit at payload byte
// w*64 + 4L (scalar.go:65-116), so the pair's word w is one 8-byte load and
// the eight threads of a row read one contiguous 64-byte super-word.
//
// Per-block semantics follow unpackUint32Scalar (scalar.go:260-309) with
// decodeUnpackHeader's validation order (unpack_block.go:127-237):
//   validate -> unpack -> exceptions (StreamVByte) -> per-lane delta
//   (zigzag or overflow-checked) -> + FOR base.
// The per-lane prefix sum over rows is a wavefront scan at thread strides
// 8/16/32; overflow detection carries a 64-bit running sum, so a wrap at
// (row, lane) shows up as a change of the carry count between row-1 and row,
// and the first wrap in v-major order is the lowest flagged flat position
// (delta.go:83-115).
#include "utl_device.h"
#include "utl_format.h"
#include "utl_kernels.h"

namespace utl {

struct BlockResult {
  uint32_t x0, x1;  // values at positions 2t, 2t+1
  int count;
  int consumed;
  bool u64single;
};

// Decodes one sub-block. forU64 selects the uint64 API's validator and
// combine/FOR64 handling (the forUint64 flag of unpackUint32Scalar).
// Returns the status word (code | overflow position << 8).
template <bool ForU64>
__device__ __forceinline__ int32_t decode_block(const uint8_t* __restrict__ buf, int64_t len, BlockResult& res) {
  const int t = lane_id();
  const int r = t >> 3;
  const int k = t & 7;
  const int p0 = 2 * t, p1 = 2 * t + 1;
  res.x0 = res.x1 = 0;
  res.count = 0;
  res.consumed = 0;
  res.u64single = false;

  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  const Hdr h = decode_header(uniform(ld_u32(buf)));
  bool combine = h.combine;
  if (ForU64) {
    if (h.int_type != 3) return UTL_ERR_UNSUPPORTED_TYPE;
  } else {
    if (h.int_type != 2 && h.int_type != 1) return UTL_ERR_UNSUPPORTED_TYPE;
    combine = false;
  }
  if ((unsigned)(h.count - 1) >= (unsigned)kBlock || (unsigned)h.bw > 32u) {
    if (h.count == 0) {
      res.consumed = kHdr;
      return UTL_OK;
    }
    if (h.count > kBlock) return UTL_ERR_INVALID_BLOCK_LENGTH;
    return UTL_ERR_INVALID_BUFFER;
  }
  const bool has_exc = h.exc > 0;
  const bool has_for = h.for_width > 0;
  const bool u64s = ForU64 && is_for64_single(h.int_type, h.for_width, combine);
  int poff = kHdr + (has_exc ? kSvbLen : 0);
  const int for_off = poff;
  if (has_for) poff += u64s ? kFor64 : for_base_bytes(h.for_width);
  if (combine) poff += kB2Len;
  const int pbytes = h.bw * 16;
  const int idx = has_exc ? exc_index_size(h.exc) : 0;
  int64_t end = (int64_t)poff + pbytes + idx;
  if (len < end) return UTL_ERR_INVALID_BUFFER;
  int svb_len = 0;
  if (has_exc) {
    svb_len = (int)uniform(ld_u16(buf + kHdr));
    if (end + svb_len > len) return UTL_ERR_INVALID_BUFFER;
  }
  uint32_t for_base = 0;
  if (has_for && !u64s) for_base = uniform(ld_bytes(buf + for_off, for_base_bytes(h.for_width)));

  // ---- UTL bit-unpack (scalar.go:65-116) ----
  uint32_t x0 = 0, x1 = 0;
  if (h.bw > 0) {
    const uint8_t* P = buf + poff;
    const uint32_t s = (uint32_t)((uintptr_t)P & 3);
    const int bitpos = r * h.bw;
    const int w = bitpos >> 5;
    const int o = bitpos & 31;
    uint32_t a0, a1;
    ld_pair(P + w * kSuperWord + 8 * k, s, a0, a1);
    uint64_t c0 = a0, c1 = a1;
    if (o + h.bw > 32) {
      uint32_t b0, b1;
      ld_pair(P + (w + 1) * kSuperWord + 8 * k, s, b0, b1);
      c0 |= (uint64_t)b0 << 32;
      c1 |= (uint64_t)b1 << 32;
    }
    const uint32_t m = low_mask(h.bw);
    x0 = (uint32_t)(c0 >> o) & m;
    x1 = (uint32_t)(c1 >> o) & m;
  }

  int consumed = poff + pbytes;
  // ---- exceptions (exceptions.go:194-284, :32-88) ----
  if (has_exc) {
    if (h.exc > h.count) return UTL_ERR_INVALID_BUFFER;
    const int ctrl = svb_ctrl_count(h.exc);
    if (ctrl > svb_len) return UTL_ERR_INVALID_BUFFER;
    const int exc_start = poff + pbytes;
    const int svb_start = exc_start + idx;
    // Exception indices i0 = 2t and i1 = 2t+1 share control byte t >> 1.
    const int i0 = 2 * t, i1 = 2 * t + 1;
    const uint32_t cb = (i0 < h.exc) ? ld_u8(buf + svb_start + (i0 >> 2)) : 0u;
    const int l0 = (i0 < h.exc) ? (int)((cb >> (2 * (i0 & 3))) & 3) + 1 : 0;
    const int l1 = (i1 < h.exc) ? (int)((cb >> (2 * (i1 & 3))) & 3) + 1 : 0;
    const uint32_t incl = wave_incl_scan((uint32_t)(l0 + l1));
    const int total = (int)uniform(__shfl(incl, 63, 64));
    // Control bytes must not describe more data than svbLen holds.
    if (ctrl + total > svb_len) return UTL_ERR_INVALID_BUFFER;
    const int d0 = svb_start + ctrl + (int)incl - (l0 + l1);
    uint32_t hb0 = l0 ? ld_bytes(buf + d0, l0) : 0u;
    uint32_t hb1 = l1 ? ld_bytes(buf + d0 + l0, l1) : 0u;
    hb0 = go_shl32(hb0, h.bw);
    hb1 = go_shl32(hb1, h.bw);
    if (h.exc <= kExcBitmap) {
      // Sorted position list. Full blocks mask positions with 0x7F,
      // partial blocks skip positions >= count (exceptions.go:211-230).
      const uint32_t q0 = (i0 < h.exc) ? ld_u8(buf + exc_start + i0) : 0u;
      const uint32_t q1 = (i1 < h.exc) ? ld_u8(buf + exc_start + i1) : 0u;
      for (int i = 0; i < h.exc; ++i) {
        const uint32_t pos = (i & 1) ? __builtin_amdgcn_readlane(q1, i >> 1) : __builtin_amdgcn_readlane(q0, i >> 1);
        const uint32_t hv = (i & 1) ? __builtin_amdgcn_readlane(hb1, i >> 1) : __builtin_amdgcn_readlane(hb0, i >> 1);
        int tgt;
        if (h.count == kBlock) {
          tgt = (int)(pos & 0x7F);
        } else {
          if ((int)pos >= h.count) continue;
          tgt = (int)pos;
        }
        if (tgt == p0) x0 |= hv;
        if (tgt == p1) x1 |= hv;
      }
    } else {
      // 128-bit bitmap masked to count; rank of a set bit = index of its
      // high bits (exceptions.go:246-284).
      uint64_t w0 = ld_u64(buf + exc_start), w1 = ld_u64(buf + exc_start + 8);
      w0 = ((uint64_t)uniform((uint32_t)(w0 >> 32)) << 32) | uniform((uint32_t)w0);
      w1 = ((uint64_t)uniform((uint32_t)(w1 >> 32)) << 32) | uniform((uint32_t)w1);
      if (h.count < kBlock) {
        if (h.count < 64) {
          w1 = 0;
          w0 &= (1ull << h.count) - 1;
        } else {
          w1 &= (1ull << (h.count - 64)) - 1;
        }
      }
      if (__popcll(w0) + __popcll(w1) <= h.exc) {
        const uint64_t w = t < 32 ? w0 : w1;
        const int bp = p0 & 63;
        const bool s0 = (w >> bp) & 1;
        const bool s1 = (w >> (bp + 1)) & 1;
        const int rank0 = (t < 32 ? 0 : __popcll(w0)) + __popcll(w & ((1ull << bp) - 1));
        const int rank1 = rank0 + (s0 ? 1 : 0);
        const uint32_t a0 = __shfl(hb0, (rank0 >> 1) & 63, 64);
        const uint32_t a1 = __shfl(hb1, (rank0 >> 1) & 63, 64);
        const uint32_t b0 = __shfl(hb0, (rank1 >> 1) & 63, 64);
        const uint32_t b1 = __shfl(hb1, (rank1 >> 1) & 63, 64);
        if (s0) x0 |= (rank0 & 1) ? a1 : a0;
        if (s1) x1 |= (rank1 & 1) ? b1 : b0;
      }
    }
    consumed = svb_start + svb_len;
  }

  // ---- per-lane delta decode (delta.go:50-115) ----
  if (h.delta) {
    if (h.zigzag) {
      x0 = zz_dec(x0);
      x1 = zz_dec(x1);
#pragma unroll
      for (int d = 8; d < 64; d <<= 1) {
        const uint32_t y0 = __shfl_up(x0, d, 64);
        const uint32_t y1 = __shfl_up(x1, d, 64);
        if (t >= d) {
          x0 += y0;
          x1 += y1;
        }
      }
    } else {
      uint64_t s0 = x0, s1 = x1;
#pragma unroll
      for (int d = 8; d < 64; d <<= 1) {
        const uint64_t y0 = __shfl_up(s0, d, 64);
        const uint64_t y1 = __shfl_up(s1, d, 64);
        if (t >= d) {
          s0 += y0;
          s1 += y1;
        }
      }
      const uint32_t c0 = (uint32_t)(s0 >> 32), c1 = (uint32_t)(s1 >> 32);
      const uint32_t pc0 = __shfl_up(c0, 8, 64), pc1 = __shfl_up(c1, 8, 64);
      const bool f0 = t >= 8 && c0 != pc0 && p0 < h.count;
      const bool f1 = t >= 8 && c1 != pc1 && p1 < h.count;
      const uint64_t m0 = __ballot(f0), m1 = __ballot(f1);
      if (m0 | m1) {
        const int q0 = m0 ? 2 * __ffsll((long long)m0) - 2 : 1 << 20;
        const int q1 = m1 ? 2 * __ffsll((long long)m1) - 1 : 1 << 20;
        return UTL_ERR_OVERFLOW | ((q0 < q1 ? q0 : q1) << 8);
      }
      x0 = (uint32_t)s0;
      x1 = (uint32_t)s1;
    }
  }
  if (has_for && !u64s) {
    x0 += for_base;
    x1 += for_base;
  }
  res.x0 = x0;
  res.x1 = x1;
  res.count = h.count;
  res.consumed = consumed;
  res.u64single = u64s;
  return UTL_OK;
}

struct Block64Result {
  uint64_t y0, y1;
  int count;
  int consumed;
};

// uint64.go:191-266 unpackUint64Block
__device__ __forceinline__ int32_t decode_block_u64(const uint8_t* __restrict__ buf, int64_t len, Block64Result& res) {
  res.y0 = res.y1 = 0;
  res.count = 0;
  res.consumed = 0;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  const Hdr h = decode_header(uniform(ld_u32(buf)));
  if (h.int_type != 3) return UTL_ERR_UNSUPPORTED_TYPE;
  BlockResult lo;
  if (!h.combine) {
    uint64_t base = 0;
    if (is_for64_single(h.int_type, h.for_width, h.combine)) {
      const int boff = kHdr + (h.exc > 0 ? kSvbLen : 0);
      if (len < boff + kFor64) return UTL_ERR_INVALID_BUFFER;
      base = ld_u64(buf + boff);
    }
    const int32_t st = decode_block<true>(buf, len, lo);
    if (st) return st;
    res.y0 = (uint64_t)lo.x0 + base;
    res.y1 = (uint64_t)lo.x1 + base;
    res.count = lo.count;
    res.consumed = lo.consumed;
    return UTL_OK;
  }
  int32_t st = decode_block<true>(buf, len, lo);
  if (st) return st;
  const bool has_exc = h.exc > 0;
  const int forb = for_base_bytes(h.for_width);
  const int meta = payload_offset(forb, has_exc, true);
  if (len < meta) return UTL_ERR_INVALID_BUFFER;
  const int b2len = (int)uniform(ld_u16(buf + kHdr + (has_exc ? kSvbLen : 0) + forb));
  const int64_t b2s = lo.consumed;
  if (b2s + b2len > len) return UTL_ERR_INVALID_BUFFER;
  BlockResult hi;
  st = decode_block<true>(buf + b2s, b2len, hi);
  if (st) return st;
  if (hi.count != lo.count) return UTL_ERR_INVALID_BUFFER;
  res.y0 = ((uint64_t)hi.x0 << 32) | lo.x0;
  res.y1 = ((uint64_t)hi.x1 << 32) | lo.x1;
  res.count = lo.count;
  res.consumed = lo.consumed + b2len;
  return UTL_OK;
}

__device__ __forceinline__ void store_u32(uint32_t* __restrict__ out, uint64_t o, int count, uint32_t x0, uint32_t x1) {
  const int t = lane_id();
  const int p0 = 2 * t;
  uint32_t* dst = out + o + p0;
  if (p0 + 1 < count && (o & 1) == 0) {
    *(uint2*)dst = make_uint2(x0, x1);
  } else {
    if (p0 < count) dst[0] = x0;
    if (p0 + 1 < count) dst[1] = x1;
  }
}

__device__ __forceinline__ void store_u64(uint64_t* __restrict__ out, uint64_t o, int count, uint64_t y0, uint64_t y1) {
  const int t = lane_id();
  const int p0 = 2 * t;
  uint64_t* dst = out + o + p0;
  if (p0 + 1 < count && (o & 1) == 0) {
    ulonglong2 v;
    v.x = y0;
    v.y = y1;
    *(ulonglong2*)dst = v;
  } else {
    if (p0 < count) dst[0] = y0;
    if (p0 + 1 < count) dst[1] = y1;
  }
}

template <bool U64>
__device__ __forceinline__ void decode_one(const uint8_t* __restrict__ src, const uint64_t* __restrict__ boff,
                                           const uint64_t* __restrict__ ooff, uint64_t b, void* __restrict__ out,
                                           int32_t* __restrict__ status, int32_t* __restrict__ consumed_out) {
  const uint64_t o0 = boff[b], o1 = boff[b + 1];
  const uint8_t* buf = src + o0;
  const int64_t len = (int64_t)(o1 - o0);
  const uint64_t out0 = ooff[b];
  const uint64_t slot = ooff[b + 1] - out0;
  int32_t st;
  int consumed;
  // A block that decodes to more values than its out_off slot disagrees with
  // the index: nothing is written and the block reports UTL_ERR_ARGUMENT.
  if (U64) {
    Block64Result res;
    st = decode_block_u64(buf, len, res);
    consumed = res.consumed;
    if (st == UTL_OK && (uint64_t)res.count > slot) st = UTL_ERR_ARGUMENT;
    if (st == UTL_OK && res.count > 0) store_u64((uint64_t*)out, out0, res.count, res.y0, res.y1);
  } else {
    BlockResult res;
    st = decode_block<false>(buf, len, res);
    consumed = res.consumed;
    if (st == UTL_OK && (uint64_t)res.count > slot) st = UTL_ERR_ARGUMENT;
    if (st == UTL_OK && res.count > 0) store_u32((uint32_t*)out, out0, res.count, res.x0, res.x1);
  }
  if (lane_id() == 0) {
    status[b] = st;
    if (consumed_out) consumed_out[b] = st == UTL_OK ? consumed : 0;
  }
}

// One wave per block, a flat grid of waves.
template <bool U64>
__global__ __launch_bounds__(256) void k_decode(const uint8_t* __restrict__ src, const uint64_t* __restrict__ boff,
                                                const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                                void* __restrict__ out, int32_t* __restrict__ status,
                                                int32_t* __restrict__ consumed_out) {
  const uint64_t b = (uint64_t)blockIdx.x * (blockDim.x >> 6) + uniform(threadIdx.x >> 6);
  if (b >= nblocks) return;
  decode_one<U64>(src, boff, ooff, b, out, status, consumed_out);
}

// Persistent form: a fixed grid of waves pulls runs of `bpw` blocks from a
// device work counter until the queue is drained. Every wave reaches the
// exit test, so the grid always drains.
template <bool U64>
__global__ __launch_bounds__(256) void k_decode_persistent(const uint8_t* __restrict__ src,
                                                           const uint64_t* __restrict__ boff,
                                                           const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                                           void* __restrict__ out, int32_t* __restrict__ status,
                                                           int32_t* __restrict__ consumed_out,
                                                           unsigned long long* __restrict__ counter, uint32_t bpw) {
  for (;;) {
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(counter, (unsigned long long)bpw);
    base = __shfl(base, 0, 64);
    if (base >= nblocks) break;
    const uint64_t stop = base + bpw < nblocks ? base + bpw : nblocks;
    for (uint64_t b = base; b < stop; ++b) decode_one<U64>(src, boff, ooff, b, out, status, consumed_out);
  }
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace utl

using namespace utl;

hipError_t utl_launch_decode(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                             void* out, int32_t* status, int32_t* consumed, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const int waves_per_wg = 4;
  const uint64_t grid = (nblocks + waves_per_wg - 1) / waves_per_wg;
  if (u64)
    k_decode<true><<<dim3((unsigned)grid), dim3(64 * waves_per_wg), 0, stream>>>(src, boff, ooff, nblocks, out, status,
                                                                                   consumed);
  else
    k_decode<false><<<dim3((unsigned)grid), dim3(64 * waves_per_wg), 0, stream>>>(src, boff, ooff, nblocks, out,
                                                                                    status, consumed);
  return hipGetLastError();
}

hipError_t utl_launch_decode_persistent(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff,
                                        uint64_t nblocks, void* out, int32_t* status, int32_t* consumed,
                                        unsigned long long* counter, const utl_tune& tune, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  k_fill_u64<<<1, 64, 0, stream>>>(counter, 0ull, 1);
  const int wpg = tune.waves_per_wg;
  dim3 block(64 * wpg), grid(tune.grid_wgs);
  if (u64)
    k_decode_persistent<true><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, out, status, consumed, counter,
                                                          tune.blocks_per_wave);
  else
    k_decode_persistent<false><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, out, status, consumed, counter,
                                                           tune.blocks_per_wave);
  return hipGetLastError();
}
