// utl_device.h -- wavefront primitives and byte readers for the gfx950 kernels.
//
// Byte readers only ever load naturally aligned dwords that contain at least
// one requested byte, so a read at the very end of a caller's buffer never
// touches a page the buffer does not own, whatever the byte alignment of the
// block (blocks are concatenated at arbitrary byte offsets).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace utl {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// n (1..4) little-endian bytes at p.
// (Pointer arithmetic on p itself, not an integer round trip, keeps the
// compiler's global address-space inference, so these stay global_load.)
__device__ __forceinline__ uint32_t ld_bytes(const uint8_t* p, int n) {
  const uint32_t s = (uint32_t)(reinterpret_cast<uintptr_t>(p) & 3);
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p - s);
  uint32_t lo = q[0];
  uint32_t hi = (s + (uint32_t)n > 4) ? q[1] : 0u;
  uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, s);
  return n >= 4 ? v : (v & ((1u << (8 * n)) - 1u));
}

__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) { return ld_bytes(p, 4); }
__device__ __forceinline__ uint32_t ld_u16(const uint8_t* p) { return ld_bytes(p, 2); }
__device__ __forceinline__ uint32_t ld_u8(const uint8_t* p) { return ld_bytes(p, 1); }
__device__ __forceinline__ uint64_t ld_u64(const uint8_t* p) {
  return (uint64_t)ld_u32(p) | ((uint64_t)ld_u32(p + 4) << 32);
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int t = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (t >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t x) {
  const int t = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if (t >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_reduce_min(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor(x, d, 64));
  return x;
}

__device__ __forceinline__ uint32_t wave_reduce_max(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = max(x, (uint32_t)__shfl_xor(x, d, 64));
  return x;
}


__device__ __forceinline__ int bit_len32(uint32_t x) { return x ? 32 - __clz((int)x) : 0; }

// ---- reductions and scans within aligned groups of 8 lanes (DPP) ----
// Every lane of a group must be active; groups never read other groups.

// value of the lane given by a DPP control: quad_perm [1,0,3,2] (0xB1),
// quad_perm [2,3,0,1] (0x4E), row_half_mirror (0x141, lane i <- lane 7-i of
// its 8 lanes)
template <int Ctrl>
__device__ __forceinline__ uint32_t g8_dpp(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, Ctrl, 0xF, 0xF, false);
}

__device__ __forceinline__ uint32_t g8_or(uint32_t x) {
  x |= g8_dpp<0xB1>(x);
  x |= g8_dpp<0x4E>(x);
  return x | g8_dpp<0x141>(x);
}
__device__ __forceinline__ uint32_t g8_add(uint32_t x) {
  x += g8_dpp<0xB1>(x);
  x += g8_dpp<0x4E>(x);
  return x + g8_dpp<0x141>(x);
}
__device__ __forceinline__ uint32_t g8_min(uint32_t x) {
  x = min(x, g8_dpp<0xB1>(x));
  x = min(x, g8_dpp<0x4E>(x));
  return min(x, g8_dpp<0x141>(x));
}
__device__ __forceinline__ uint32_t g8_max(uint32_t x) {
  x = max(x, g8_dpp<0xB1>(x));
  x = max(x, g8_dpp<0x4E>(x));
  return max(x, g8_dpp<0x141>(x));
}

// Inclusive prefix sum over the group (lane j = lane & 7): row_shr:1,2,4
// within 16-lane rows, keeping only sources of the same group.
__device__ __forceinline__ uint32_t g8_incl_scan(uint32_t x) {
  const int j = lane_id() & 7;
  uint32_t y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);
  x += j >= 1 ? y : 0u;
  y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);
  x += j >= 2 ? y : 0u;
  y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);
  x += j >= 4 ? y : 0u;
  return x;
}

// Inclusive scan over the wave's eight 8-lane groups of a value that is equal
// within each group (DPP row_shr:8, row_bcast:15, row_bcast:31). Must run
// with all 64 lanes active.
__device__ __forceinline__ uint32_t wave_group_scan(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
  return x;
}

}  // namespace utl
