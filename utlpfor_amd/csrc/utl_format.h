// utl_format.h -- UTL-PFOR wire-format constants and header decoding shared by
// the HIP kernels and the host side of the C-ABI (__host__ __device__).
//
// Wire format (header.go:8-87, utlpfor.ksy:18-67):
//   [hdr u32][svbLen u16 if exc][FOR base 0/1/2/4, 8 for FOR64][block2Len u16
//   if combine][payload 16*bw B][exception index][StreamVByte]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/utlpfor_gpu.h"

namespace utl {

constexpr int kBlock = 128;      // values per block (header.go:9)
constexpr int kLanes = 16;       // UTL lanes (header.go:10)
constexpr int kRows = 8;         // values per lane (header.go:11)
constexpr int kSuperWord = 64;   // bytes per super-word (header.go:12)
constexpr int kHdr = 4;          // header bytes
constexpr int kSvbLen = 2;       // svbLen field bytes
constexpr int kB2Len = 2;        // block2Len field bytes
constexpr int kFor64 = 8;        // FOR64 base bytes (uint64.go:18)
constexpr int kExcBitmap = 16;   // position list up to 16, else bitmap (header.go:68)

constexpr uint32_t kTypeU16 = 1u << 13;
constexpr uint32_t kTypeU32 = 2u << 13;
constexpr uint32_t kTypeU64 = 3u << 13;
constexpr uint32_t kHSpecial = 1u << 17;
constexpr uint32_t kHReserved = 1u << 18;
constexpr uint32_t kHCombine = 1u << 19;
constexpr uint32_t kHBlock256 = 1u << 20;
constexpr uint32_t kHDelta = 1u << 22;
constexpr uint32_t kHZigZag = 1u << 23;

struct Hdr {
  int count, bw, int_type, exc, for_width;
  bool delta, zigzag, special, combine;
};

// header.go:116 decodeHeader
__host__ __device__ inline Hdr decode_header(uint32_t h) {
  Hdr d;
  d.count = (int)(h & 0xFF);
  d.bw = (int)((h >> 8) & 0x1F) * 4;
  d.int_type = (int)((h >> 13) & 3);
  d.for_width = (int)((h >> 15) & 3);
  d.exc = (int)(h >> 24);
  d.delta = (h & kHDelta) != 0;
  d.zigzag = (h & kHZigZag) != 0;
  d.special = (h & kHSpecial) != 0;
  d.combine = (h & kHCombine) != 0;
  return d;
}

// header.go:102 encodeHeader
__host__ __device__ inline uint32_t encode_header(int count, int bw, int exc, uint32_t flags) {
  return (uint32_t)(count & 0xFF) | ((uint32_t)((bw >> 2) & 0x1F) << 8) | flags |
         ((uint32_t)(exc & 0xFF) << 24);
}

__host__ __device__ inline int payload_bytes(int bw) { return ((bw + 3) >> 2) << 6; }  // header.go:93
__host__ __device__ inline int for_base_bytes(int fw) { return (1 << fw) >> 1; }        // header.go:97
__host__ __device__ inline int exc_index_size(int exc) { return exc <= kExcBitmap ? exc : 16; }
__host__ __device__ inline int svb_ctrl_count(int n) { return (n + 3) >> 2; }            // svb.go:4
// svb.go:9-24: data bytes described by one control byte
__host__ __device__ inline int svb_ctrl_block_size(uint32_t c) {
  return 4 + (c & 3) + ((c >> 2) & 3) + ((c >> 4) & 3) + ((c >> 6) & 3);
}
// uint64.go:70 isFor64SingleBlock
__host__ __device__ inline bool is_for64_single(int t, int fw, bool combine) {
  return t == 3 && !combine && fw == 3;
}
// uint64.go:77 forBaseBytesForBlock
__host__ __device__ inline int for_base_bytes_for_block(int fw, int t, bool combine) {
  return is_for64_single(t, fw, combine) ? kFor64 : for_base_bytes(fw);
}
// header.go:157 payloadOffset
__host__ __device__ inline int payload_offset(int forb, bool has_exc, bool combine) {
  return kHdr + (has_exc ? kSvbLen : 0) + forb + (combine ? kB2Len : 0);
}
// Go's shift semantics: a count >= the width gives 0.
__host__ __device__ inline uint32_t go_shl32(uint32_t x, int s) { return s >= 32 ? 0u : (x << s); }
__host__ __device__ inline uint32_t low_mask(int bw) { return bw >= 32 ? 0xFFFFFFFFu : ((1u << bw) - 1u); }
__host__ __device__ inline uint32_t zz_enc(uint32_t v) { return (v << 1) ^ (uint32_t)((int32_t)v >> 31); }
__host__ __device__ inline uint32_t zz_dec(uint32_t v) { return (v >> 1) ^ (0u - (v & 1u)); }

// Host-side little-endian readers for the per-call header functions.
inline uint32_t host_ld16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
inline uint32_t host_ld32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// block_length.go:8 BlockLength over a byte reader (host or device).
// rd16/rd32(off) return the little-endian field at byte offset off.
template <class Rd32, class Rd16>
__host__ __device__ inline int block_length_impl(int64_t len, Rd32 rd32, Rd16 rd16, int64_t* out) {
  *out = 0;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  uint32_t header = rd32(0);
  Hdr h = decode_header(header);
  if (h.count > kBlock) return UTL_ERR_INVALID_BLOCK_LENGTH;
  if (h.bw > 32) return UTL_ERR_INVALID_BUFFER;
  bool combine = h.combine && h.int_type == 3;
  if (h.int_type == 3) {
    if (header & (kHReserved | kHBlock256)) return UTL_ERR_INVALID_FLAGS;
  } else if (header & (kHReserved | kHCombine | kHBlock256)) {
    return UTL_ERR_INVALID_FLAGS;
  }
  bool has_exc = h.exc > 0;
  int forb = for_base_bytes_for_block(h.for_width, h.int_type, combine);
  int poff = payload_offset(forb, has_exc, combine);
  int64_t b1 = poff + (h.bw << 4);
  if (has_exc) {
    if (len < kHdr + kSvbLen) return UTL_ERR_INVALID_BUFFER;
    b1 += exc_index_size(h.exc) + (int64_t)rd16(kHdr);
  }
  if (!combine) {
    *out = b1;
    return UTL_OK;
  }
  if (len < poff) return UTL_ERR_INVALID_BUFFER;
  *out = b1 + (int64_t)rd16(kHdr + (has_exc ? kSvbLen : 0) + forb);
  return UTL_OK;
}

}  // namespace utl
