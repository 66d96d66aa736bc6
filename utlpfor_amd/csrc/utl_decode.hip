// utl_decode.hip -- gfx950 decode kernels for UTL-PFOR blocks.
//
// A block is decoded by 16 threads (a quarter of a wavefront, one DPP row);
// thread L owns UTL lane L, i.e. the 8 values at positions 16r + L. A
// wavefront decodes 4 blocks side by side and a 256-thread workgroup a tile
// of 16 consecutive blocks. Tiles are assigned XCD-major (workgroup g runs
// on XCD g % 8), so each XCD's L2 sees a contiguous range of the stream and
// cache lines shared by neighbouring blocks are fetched from HBM once.
//
// Each block is first staged into LDS with 16-byte loads of the aligned
// chunks that contain its bytes (one round trip to HBM), then decoded from
// LDS following unpackUint32Scalar (scalar.go:260-309) and its validation
// order (unpack_block.go:127-237):
//   validate -> unpack (one specialised path per bit width) -> exceptions
//   (StreamVByte into LDS; bitmap patch by popcount rank, position-list
//   patch by LDS OR) -> per-lane delta in registers (zigzag, or overflow
//   checked with the first wrap in v-major order) -> + FOR base.
#include "utl_device.h"
#include "utl_format.h"
#include "utl_kernels.h"

namespace utl {

constexpr int kQuad = 16;           // threads per block
constexpr int kWaveBlocks = 4;      // blocks per wavefront
constexpr int kWgWaves = 4;         // wavefronts per workgroup
constexpr int kTileBlocks = kWaveBlocks * kWgWaves;
// Bytes of a sub-block the decoder can read: header, svbLen, FOR64 base,
// block2Len (16) + payload (512) + bitmap (16) + control (32) + data (512).
constexpr int kSubExtent = 1088;
constexpr int kSlotBytes = 1120;    // >= 15 + kSubExtent + 4, multiple of 16

struct QuadLds {
  uint32_t slot[kSlotBytes / 4];    // staged sub-block bytes (at byte offset s)
  uint32_t aux[kBlock];             // exception high bits, or values (position list)
};

__device__ __forceinline__ void quad_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP row (16-lane) primitives. row_shr:n = 0x110 + n, row_ror:n = 0x120 + n.
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, Ctrl, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t quad_incl_scan(uint32_t x) {
  x += dpp<0x111>(x);
  x += dpp<0x112>(x);
  x += dpp<0x114>(x);
  x += dpp<0x118>(x);
  return x;
}

__device__ __forceinline__ uint32_t quad_min(uint32_t x) {
  x = min(x, dpp<0x121>(x));
  x = min(x, dpp<0x122>(x));
  x = min(x, dpp<0x124>(x));
  x = min(x, dpp<0x128>(x));
  return x;
}

// Value of lane 15 of this thread's row.
__device__ __forceinline__ uint32_t quad_last(uint32_t x) { return (uint32_t)__shfl((int)x, (lane_id() & ~15) | 15, 64); }

// n (1..4) little-endian bytes at byte offset off of an LDS slot.
__device__ __forceinline__ uint32_t lds_bytes(const uint32_t* slot, int off, int n) {
  const int q = off >> 2, sh = off & 3;
  const uint32_t lo = slot[q], hi = slot[q + 1];
  const uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, sh);
  return n >= 4 ? v : (v & ((1u << (8 * n)) - 1u));
}
__device__ __forceinline__ uint32_t lds_u32(const uint32_t* slot, int off) { return lds_bytes(slot, off, 4); }

// Stages nbytes at g into the slot: every 16-byte aligned chunk that holds
// one of the bytes is copied, so the block starts at slot byte (g & 15).
__device__ __forceinline__ int stage(const uint8_t* __restrict__ g, int64_t nbytes, uint32_t* slot, int L) {
  const int s = (int)(reinterpret_cast<uintptr_t>(g) & 15);
  const uint4* base = reinterpret_cast<const uint4*>(g - s);
  const int nchunks = nbytes > 0 ? (int)((s + nbytes + 15) >> 4) : 0;
  for (int c = L; c < nchunks; c += kQuad) reinterpret_cast<uint4*>(slot)[c] = base[c];
  quad_fence();
  return s;
}

// UTL unpack of lane L for a fixed bit width (scalar.go:65-116): word w of
// the lane is at payload byte 64w + 4L; value r occupies bits [r*BW, +BW).
template <int BW>
__device__ __forceinline__ void unpack_bw(const uint32_t* slot, int pbyte, int L, uint32_t* v) {
  if constexpr (BW == 0) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) v[r] = 0;
  } else {
    constexpr int NW = BW / 4;
    uint32_t w[NW + 1];
    const int base = pbyte + 4 * L;
    const int q = base >> 2, sh = base & 3;
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = __builtin_amdgcn_alignbyte(slot[q + 16 * k + 1], slot[q + 16 * k], sh);
    w[NW] = 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int bit = r * BW, wi = bit >> 5, o = bit & 31;
      if constexpr (BW == 32) {
        v[r] = w[wi];
      } else {
        constexpr uint32_t m = (1u << BW) - 1u;
        if (o + BW <= 32) v[r] = (w[wi] >> o) & m;
        else v[r] = __builtin_amdgcn_alignbit(w[wi + 1], w[wi], o) & m;
      }
    }
  }
}

__device__ __forceinline__ void unpack_any(int bw, const uint32_t* slot, int pbyte, int L, uint32_t* v) {
  switch (bw) {
    case 0: unpack_bw<0>(slot, pbyte, L, v); break;
    case 4: unpack_bw<4>(slot, pbyte, L, v); break;
    case 8: unpack_bw<8>(slot, pbyte, L, v); break;
    case 12: unpack_bw<12>(slot, pbyte, L, v); break;
    case 16: unpack_bw<16>(slot, pbyte, L, v); break;
    case 20: unpack_bw<20>(slot, pbyte, L, v); break;
    case 24: unpack_bw<24>(slot, pbyte, L, v); break;
    case 28: unpack_bw<28>(slot, pbyte, L, v); break;
    default: unpack_bw<32>(slot, pbyte, L, v); break;
  }
}

// Data bytes described by control byte c when only its first n slots count.
__device__ __forceinline__ int ctrl_bytes(uint32_t c, int n) {
  int s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) s += (k < n) ? (int)((c >> (2 * k)) & 3) + 1 : 0;
  return s;
}

struct SubResult {
  int count;
  int consumed;
  bool u64single;
};

// Decodes one staged sub-block (at slot byte s, true length len) for lane L.
// ForU64 is the uint64 API's validator and FOR64/combine handling. Returns
// the status word (code | overflow position << 8).
template <bool ForU64>
__device__ int32_t decode_sub(QuadLds* Q, int s, int64_t len, int L, uint32_t* v, SubResult& res) {
  const uint32_t* slot = Q->slot;
  res.count = 0;
  res.consumed = 0;
  res.u64single = false;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  const Hdr h = decode_header(lds_u32(slot, s));
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
  const int64_t end = (int64_t)poff + pbytes + idx;
  if (len < end) return UTL_ERR_INVALID_BUFFER;
  int svb_len = 0;
  if (has_exc) {
    svb_len = (int)lds_bytes(slot, s + kHdr, 2);
    if (end + svb_len > len) return UTL_ERR_INVALID_BUFFER;
  }
  const uint32_t for_base = (has_for && !u64s) ? lds_bytes(slot, s + for_off, for_base_bytes(h.for_width)) : 0u;

  unpack_any(h.bw, slot, s + poff, L, v);

  int consumed = poff + pbytes;
  if (has_exc) {
    // ---- exceptions (exceptions.go:32-88, :194-284) ----
    if (h.exc > h.count) return UTL_ERR_INVALID_BUFFER;
    const int ctrl = svb_ctrl_count(h.exc);
    if (ctrl > svb_len) return UTL_ERR_INVALID_BUFFER;
    const int exc_start = poff + pbytes;
    const int svb_start = exc_start + idx;
    // Thread L owns control bytes 2L, 2L+1, i.e. exceptions 8L .. 8L+7.
    const int j0 = 2 * L, j1 = 2 * L + 1;
    const uint32_t cb0 = j0 < ctrl ? lds_bytes(slot, s + svb_start + j0, 1) : 0u;
    const uint32_t cb1 = j1 < ctrl ? lds_bytes(slot, s + svb_start + j1, 1) : 0u;
    const int n0 = min(max(h.exc - 4 * j0, 0), 4), n1 = min(max(h.exc - 4 * j1, 0), 4);
    const int part = ctrl_bytes(cb0, n0) + ctrl_bytes(cb1, n1);
    const uint32_t incl = quad_incl_scan((uint32_t)part);
    const int total = (int)quad_last(incl);
    if (ctrl + total > svb_len) return UTL_ERR_INVALID_BUFFER;
    // StreamVByte decode of this thread's exceptions into aux[].
    int off = s + svb_start + ctrl + (int)incl - part;
    const int nmine = n0 + n1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k < nmine) {
        const uint32_t c = k < 4 ? (cb0 >> (2 * k)) : (cb1 >> (2 * (k - 4)));
        const int l = (int)(c & 3) + 1;
        Q->aux[8 * L + k] = go_shl32(lds_bytes(slot, off, l), h.bw);
        off += l;
      }
    }
    quad_fence();
    if (h.exc <= kExcBitmap) {
      // Sorted position list: OR each exception into its position in LDS.
      // Full blocks mask positions with 0x7F, partial blocks skip positions
      // >= count (exceptions.go:211-230).
      const uint32_t hv = L < h.exc ? Q->aux[L] : 0u;
      quad_fence();
#pragma unroll
      for (int r = 0; r < kRows; ++r) Q->aux[16 * r + L] = v[r];
      quad_fence();
      if (L < h.exc) {
        const uint32_t p = lds_bytes(slot, s + exc_start + L, 1);
        if (h.count == kBlock) atomicOr(&Q->aux[p & 0x7F], hv);
        else if ((int)p < h.count) atomicOr(&Q->aux[p], hv);
      }
      quad_fence();
#pragma unroll
      for (int r = 0; r < kRows; ++r) v[r] = Q->aux[16 * r + L];
    } else {
      // 128-bit bitmap masked to count; the rank of a set bit is the index
      // of its high bits. popcount(bitmap) > excCount applies nothing.
      uint32_t bm[4];
#pragma unroll
      for (int d = 0; d < 4; ++d) bm[d] = lds_u32(slot, s + exc_start + 4 * d);
      if (h.count < kBlock) {
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int keep = h.count - 32 * d;
          bm[d] = keep >= 32 ? bm[d] : (keep <= 0 ? 0u : (bm[d] & ((1u << keep) - 1u)));
        }
      }
      const uint32_t pre1 = __popc(bm[0]), pre2 = pre1 + __popc(bm[1]), pre3 = pre2 + __popc(bm[2]);
      if ((int)(pre3 + __popc(bm[3])) <= h.exc) {
        const uint32_t m_lo = (1u << L) - 1u, m_hi = (1u << (L + 16)) - 1u;
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          const int d = r >> 1;
          const uint32_t word = bm[d];
          const uint32_t pre = d == 0 ? 0u : d == 1 ? pre1 : d == 2 ? pre2 : pre3;
          const int bit = L + 16 * (r & 1);
          if ((word >> bit) & 1u) {
            const uint32_t rank = pre + __popc(word & ((r & 1) ? m_hi : m_lo));
            v[r] |= Q->aux[rank];
          }
        }
      }
    }
    consumed = svb_start + svb_len;
  }

  // ---- per-lane delta decode (delta.go:50-115) ----
  if (h.delta) {
    if (h.zigzag) {
      v[0] = zz_dec(v[0]);
#pragma unroll
      for (int r = 1; r < kRows; ++r) v[r] = v[r - 1] + zz_dec(v[r]);
    } else {
      uint32_t first = 0xFFFFFFFFu;
#pragma unroll
      for (int r = 1; r < kRows; ++r) {
        const uint32_t nx = v[r - 1] + v[r];
        const int pos = 16 * r + L;
        if (nx < v[r - 1] && pos < h.count && first == 0xFFFFFFFFu) first = (uint32_t)pos;
        v[r] = nx;
      }
      // v-major order is the flat position: the lowest flagged position wins
      const uint32_t m = quad_min(first);
      if (m != 0xFFFFFFFFu) return UTL_ERR_OVERFLOW | (int32_t)(m << 8);
    }
  }
  if (has_for && !u64s) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) v[r] += for_base;
  }
  res.count = h.count;
  res.consumed = consumed;
  res.u64single = u64s;
  return UTL_OK;
}

__device__ __forceinline__ int64_t cap_extent(int64_t len) { return len < kSubExtent ? len : kSubExtent; }

// Decodes block b (one quarter-wave) and stores its values.
template <bool U64>
__device__ __forceinline__ void decode_block(QuadLds* Q, const uint8_t* __restrict__ src, uint64_t o0, uint64_t o1,
                                             uint64_t out0, uint64_t slot_n, void* __restrict__ out, int32_t& st,
                                             int& consumed) {
  const int L = lane_id() & 15;
  const uint8_t* buf = src + o0;
  const int64_t len = (int64_t)(o1 - o0);
  uint32_t lo[kRows];
  SubResult r1;
  int s = stage(buf, cap_extent(len), Q->slot, L);
  if (!U64) {
    st = decode_sub<false>(Q, s, len, L, lo, r1);
    consumed = r1.consumed;
    if (st == UTL_OK && (uint64_t)r1.count > slot_n) st = UTL_ERR_ARGUMENT;
    if (st == UTL_OK) {
      uint32_t* o = (uint32_t*)out + out0;
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if (16 * r + L < r1.count) o[16 * r + L] = lo[r];
    }
    return;
  }
  // uint64.go:191-266 unpackUint64Block
  consumed = 0;
  if (len < kHdr) {
    st = UTL_ERR_INVALID_BUFFER;
    return;
  }
  const Hdr h = decode_header(lds_u32(Q->slot, s));
  if (h.int_type != 3) {
    st = UTL_ERR_UNSUPPORTED_TYPE;
    return;
  }
  uint64_t y[kRows];
  int count;
  if (!h.combine) {
    uint64_t base = 0;
    if (is_for64_single(h.int_type, h.for_width, h.combine)) {
      const int boff = kHdr + (h.exc > 0 ? kSvbLen : 0);
      if (len < boff + kFor64) {
        st = UTL_ERR_INVALID_BUFFER;
        return;
      }
      base = (uint64_t)lds_u32(Q->slot, s + boff) | ((uint64_t)lds_u32(Q->slot, s + boff + 4) << 32);
    }
    st = decode_sub<true>(Q, s, len, L, lo, r1);
    if (st) return;
#pragma unroll
    for (int r = 0; r < kRows; ++r) y[r] = (uint64_t)lo[r] + base;
    count = r1.count;
    consumed = r1.consumed;
  } else {
    st = decode_sub<true>(Q, s, len, L, lo, r1);
    if (st) return;
    const bool has_exc = h.exc > 0;
    const int forb = for_base_bytes(h.for_width);
    const int meta = payload_offset(forb, has_exc, true);
    if (len < meta) {
      st = UTL_ERR_INVALID_BUFFER;
      return;
    }
    const int b2len = (int)lds_bytes(Q->slot, s + kHdr + (has_exc ? kSvbLen : 0) + forb, 2);
    const int64_t b2s = r1.consumed;
    if (b2s + b2len > len) {
      st = UTL_ERR_INVALID_BUFFER;
      return;
    }
    quad_fence();
    uint32_t hi[kRows];
    SubResult r2;
    s = stage(buf + b2s, cap_extent(b2len), Q->slot, L);
    st = decode_sub<true>(Q, s, b2len, L, hi, r2);
    if (st) return;
    if (r2.count != r1.count) {
      st = UTL_ERR_INVALID_BUFFER;
      return;
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r) y[r] = ((uint64_t)hi[r] << 32) | lo[r];
    count = r1.count;
    consumed = r1.consumed + b2len;
  }
  if ((uint64_t)count > slot_n) {
    st = UTL_ERR_ARGUMENT;
    return;
  }
  uint64_t* o = (uint64_t*)out + out0;
#pragma unroll
  for (int r = 0; r < kRows; ++r)
    if (16 * r + L < count) o[16 * r + L] = y[r];
}

// Workgroup g runs on XCD g % 8; tile(g) makes each XCD's workgroups cover
// one contiguous range of tiles.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t g, uint64_t per_xcd) { return (g & 7) * per_xcd + (g >> 3); }

template <bool U64>
__device__ __forceinline__ void decode_tile_block(QuadLds* Q, const uint8_t* __restrict__ src,
                                                  const uint64_t* __restrict__ boff, const uint64_t* __restrict__ ooff,
                                                  uint64_t b, void* __restrict__ out, int32_t* __restrict__ status,
                                                  int32_t* __restrict__ consumed_out) {
  const uint64_t o0 = boff[b], o1 = boff[b + 1];
  const uint64_t out0 = ooff[b], out1 = ooff[b + 1];
  int32_t st;
  int consumed;
  decode_block<U64>(Q, src, o0, o1, out0, out1 - out0, out, st, consumed);
  if ((lane_id() & 15) == 0) {
    status[b] = st;
    if (consumed_out) consumed_out[b] = st == UTL_OK ? consumed : 0;
  }
}

// Flat grid: one tile of 16 blocks per workgroup, XCD-major.
template <bool U64>
__global__ __launch_bounds__(256) void k_decode(const uint8_t* __restrict__ src, const uint64_t* __restrict__ boff,
                                                const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                                void* __restrict__ out, int32_t* __restrict__ status,
                                                int32_t* __restrict__ consumed_out, uint64_t per_xcd) {
  __shared__ QuadLds lds[kTileBlocks];
  const uint64_t tile = xcd_tile(blockIdx.x, per_xcd);
  const int quad = (int)(threadIdx.x >> 4);
  const uint64_t b = tile * kTileBlocks + quad;
  if (b >= nblocks) return;  // whole quarter exits together
  decode_tile_block<U64>(&lds[quad], src, boff, ooff, b, out, status, consumed_out);
}

// Persistent form: a fixed grid; workgroups on XCD x take tiles from XCD x's
// own work counter (counter[x]) over XCD x's contiguous range of tiles, so
// load balance stays inside the XCD. Every workgroup reaches the exit test.
template <bool U64>
__global__ __launch_bounds__(256) void k_decode_persistent(const uint8_t* __restrict__ src,
                                                           const uint64_t* __restrict__ boff,
                                                           const uint64_t* __restrict__ ooff, uint64_t nblocks,
                                                           void* __restrict__ out, int32_t* __restrict__ status,
                                                           int32_t* __restrict__ consumed_out,
                                                           unsigned long long* __restrict__ counter, uint64_t per_xcd,
                                                           uint32_t tiles_per_grab) {
  __shared__ QuadLds lds[kTileBlocks];
  __shared__ unsigned long long grab;
  const uint32_t xcd = blockIdx.x & 7;
  const int quad = (int)(threadIdx.x >> 4);
  const uint64_t ntiles = (nblocks + kTileBlocks - 1) / kTileBlocks;
  const uint64_t first = xcd * per_xcd;
  const uint64_t last = min(first + per_xcd, ntiles);
  for (;;) {
    if (threadIdx.x == 0) grab = atomicAdd(&counter[xcd * 8], (unsigned long long)tiles_per_grab);
    __syncthreads();
    const uint64_t t0 = first + grab;
    __syncthreads();
    if (t0 >= last) break;
    const uint64_t t1 = min(t0 + tiles_per_grab, last);
    for (uint64_t tile = t0; tile < t1; ++tile) {
      const uint64_t b = tile * kTileBlocks + quad;
      if (b < nblocks) decode_tile_block<U64>(&lds[quad], src, boff, ooff, b, out, status, consumed_out);
      quad_fence();
    }
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
  const uint64_t ntiles = (nblocks + kTileBlocks - 1) / kTileBlocks;
  const uint64_t per_xcd = (ntiles + 7) / 8;
  const dim3 grid((unsigned)(per_xcd * 8)), block(64 * kWgWaves);
  if (u64)
    k_decode<true><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, out, status, consumed, per_xcd);
  else
    k_decode<false><<<grid, block, 0, stream>>>(src, boff, ooff, nblocks, out, status, consumed, per_xcd);
  return hipGetLastError();
}

hipError_t utl_launch_decode_persistent(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff,
                                        uint64_t nblocks, void* out, int32_t* status, int32_t* consumed,
                                        unsigned long long* counter, const utl_tune& tune, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  k_fill_u64<<<1, 64, 0, stream>>>(counter, 0ull, 64);
  const uint64_t ntiles = (nblocks + kTileBlocks - 1) / kTileBlocks;
  const uint64_t per_xcd = (ntiles + 7) / 8;
  const unsigned grid = (unsigned)((tune.grid_wgs + 7) / 8 * 8);
  const uint32_t grab = tune.blocks_per_wave ? tune.blocks_per_wave : 1;
  if (u64)
    k_decode_persistent<true><<<dim3(grid), dim3(256), 0, stream>>>(src, boff, ooff, nblocks, out, status, consumed,
                                                                     counter, per_xcd, grab);
  else
    k_decode_persistent<false><<<dim3(grid), dim3(256), 0, stream>>>(src, boff, ooff, nblocks, out, status, consumed,
                                                                      counter, per_xcd, grab);
  return hipGetLastError();
}
