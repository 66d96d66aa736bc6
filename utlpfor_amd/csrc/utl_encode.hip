// utl_encode.hip -- gfx950 encode kernels for UTL-PFOR blocks.
//
// One wavefront encodes one block. Thread t owns input positions 2t and
// 2t+1 (row t >> 3, lanes 2(t & 7) and 2(t & 7)+1). The block is assembled
// byte by byte in a per-wave LDS buffer and then written out with coalesced
// byte stores into a fixed-size slot; a scan of the block lengths and a
// compaction pass produce the concatenated stream.
//
// The steps follow packBlockScalar (scalar.go:126-235) exactly:
//   FOR decision decideFORFull (for.go:95-118) from wave min/max and the
//   8 threshold ballots of the raw values; per-lane delta with zigzag of all
//   values on any decrease (delta.go:23-47); bit width by the integer cost
//   model over exception counts (select_bitwidth.go:41-121; ties keep the
//   wider width); payload words gathered from LDS (scalar.go:7-59);
//   exceptions compacted by ballot rank into a sorted position list or a
//   bitmap (exceptions.go:138-186) with their high bits StreamVByte-encoded
//   ("1234" codes, control bytes first).
// uint64 blocks follow packUint64Scalar (uint64.go:302-333): FOR64, fit-in-32
// or two sub-blocks with block2Len patched into the first.
#include "utl_device.h"
#include "utl_encode.h"
#include "utl_format.h"

namespace utl {

constexpr int kEncWaves = 4;        // wavefronts per workgroup
constexpr int kEncBuf = 2112;       // per-wave LDS block buffer (>= 2038)

struct EncLds {
  uint8_t blk[kEncBuf];
  uint32_t vals[kBlock];
  uint8_t codes[kBlock];
};

// select_bitwidth.go:56-72 chooseBestFromHist, fed with exc[si] = number of
// values whose step index exceeds si (equivalent to the suffix sums).
__device__ __forceinline__ void choose_best(const int* exc, int max_si, int& best_w, int& best_exc) {
  const int bpv[9] = {1, 1, 1, 2, 2, 3, 3, 4, 4};
  int cost = kHdr + (max_si << 6);
  best_w = max_si << 2;
  best_exc = 0;
  for (int si = max_si - 1; si >= 0; --si) {
    const int e = exc[si];
    const int c = kHdr + (si << 6) + kSvbLen + (e < kExcBitmap ? e : kExcBitmap) + ((e + 3) >> 2) + e * bpv[max_si - si];
    if (c < cost) {
      cost = c;
      best_w = si << 2;
      best_exc = e;
    }
  }
}

// Counts of valid values above each step threshold 2^(4si)-1, si = 0..7.
__device__ __forceinline__ void threshold_counts(uint32_t x0, uint32_t x1, bool v0, bool v1, int* exc) {
#pragma unroll
  for (int si = 0; si < 8; ++si) {
    const uint32_t thr = (si == 0) ? 0u : ((1u << (4 * si)) - 1u);
    exc[si] = __popcll(__ballot(v0 && x0 > thr)) + __popcll(__ballot(v1 && x1 > thr));
  }
  exc[8] = 0;
}

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void put_bytes(uint8_t* p, uint64_t v, int n) {
  for (int i = 0; i < n; ++i) p[i] = (uint8_t)(v >> (8 * i));
}

__device__ __forceinline__ int step_index_len(uint32_t x) { return (bit_len32(x) + 3) >> 2; }

// Packs one sub-block whose work values (positions >= n are 0) are x0/x1
// into L->blk + at. Writes the header (with type_flags), svbLen, FOR base
// (or the FOR64 base when for64 is set, the inner block being NoFOR), leaves
// the 2-byte block2Len slot of a combine block for the caller, and returns
// the sub-block length. x0/x1 return the work values after FOR/delta/zigzag
// (the reference's in-place result).
__device__ int pack_sub_block(EncLds* L, int at, uint32_t flag, uint32_t& x0, uint32_t& x1, int n,
                              uint32_t type_flags, bool combine, bool for64, uint64_t base64) {
  const int t = lane_id();
  const int p0 = 2 * t, p1 = 2 * t + 1;
  const bool v0 = p0 < n, v1 = p1 < n;
  uint32_t hflags = type_flags | ((flag & UTL_SPECIAL) << 14);

  // ---- FOR (for.go:95-118) ----
  bool use_for = false;
  uint32_t base = 0;
  int fw = 0;
  if (!(flag & UTL_NOFOR)) {
    const uint32_t mn = uniform(wave_reduce_min(min(v0 ? x0 : 0xFFFFFFFFu, v1 ? x1 : 0xFFFFFFFFu)));
    const uint32_t mx = uniform(wave_reduce_max(max(v0 ? x0 : 0u, v1 ? x1 : 0u)));
    if (mn != 0) {
      const int for_max = ((bit_len32(mx - mn) + 3) >> 2) << 2;
      const int raw_max = ((bit_len32(mx) + 3) >> 2) << 2;
      if (for_max < raw_max) {
        int exc[9];
        threshold_counts(x0, x1, v0, v1, exc);
        int std_w, dummy;
        choose_best(exc, step_index_len(mx), std_w, dummy);
        if (for_max < std_w) {
          use_for = true;
          base = mn;
          const int bytes = (bit_len32(mn) + 7) >> 3;
          fw = bytes == 0 ? 0 : bytes == 1 ? 1 : bytes == 2 ? 2 : 3;
          if (v0) x0 -= base;
          if (v1) x1 -= base;
          hflags |= (uint32_t)fw << 15;
        }
      }
    }
  }
  // ---- per-lane delta (delta.go:23-47) ----
  if (flag & UTL_DELTA) {
    const uint32_t y0 = __shfl_up(x0, 8, 64), y1 = __shfl_up(x1, 8, 64);
    const bool d0 = t >= 8 && v0, d1 = t >= 8 && v1;
    const bool zz = (__ballot((d0 && x0 < y0) || (d1 && x1 < y1))) != 0;
    if (d0) x0 -= y0;
    if (d1) x1 -= y1;
    if (zz) {
      if (v0) x0 = zz_enc(x0);
      if (v1) x1 = zz_enc(x1);
      hflags |= kHZigZag;
    }
    hflags |= kHDelta;
  }
  // ---- bit width (select_bitwidth.go:20-121) ----
  int bw, exc = 0;
  const uint32_t orall = uniform(wave_reduce_or((v0 ? x0 : 0u) | (v1 ? x1 : 0u)));
  const int max_si = step_index_len(orall);
  if (flag & UTL_NOPATCH) {
    bw = max_si << 2;
  } else {
    int cnt[9];
    threshold_counts(x0, x1, v0, v1, cnt);
    choose_best(cnt, max_si, bw, exc);
  }
  const bool has_exc = exc > 0;
  const int forb = for64 ? kFor64 : for_base_bytes(fw);
  const int poff = kHdr + (has_exc ? kSvbLen : 0) + forb + (combine ? kB2Len : 0);
  const int pbytes = bw * 16;
  uint8_t* B = L->blk + at;

  // ---- payload (scalar.go:7-59) ----
  const uint32_t m = low_mask(bw);
  L->vals[p0] = v0 ? (x0 & m) : 0u;
  L->vals[p1] = v1 ? (x1 & m) : 0u;
  lds_fence();
  const int ndw = 4 * bw;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int j = t + 64 * half;
    if (j < ndw) {
      const int w = j >> 4, lane = j & 15;
      uint32_t acc = 0;
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int s = r * bw - 32 * w;
        if (s > -bw && s < 32) {
          const uint32_t v = L->vals[16 * r + lane];
          acc |= s >= 0 ? (v << s) : (v >> (-s));
        }
      }
      put_bytes(B + poff + 4 * j, acc, 4);
    }
  }
  // ---- header, FOR base ----
  if (t == 0) {
    put_bytes(B, encode_header(n, bw, exc, hflags | (for64 ? (3u << 15) : 0u)), 4);
    const int boff = kHdr + (has_exc ? kSvbLen : 0);
    if (for64) put_bytes(B + boff, base64, 8);
    else if (use_for) put_bytes(B + boff, base, for_base_bytes(fw));
  }
  int len = poff + pbytes;
  // ---- exceptions (exceptions.go:138-186) ----
  if (has_exc) {
    const bool e0 = v0 && x0 > m, e1 = v1 && x1 > m;
    const uint64_t b0 = __ballot(e0), b1 = __ballot(e1);
    const uint64_t below = (t == 0) ? 0ull : (~0ull >> (64 - t));
    const int rank0 = __popcll(b0 & below) + __popcll(b1 & below);
    const int rank1 = rank0 + (e0 ? 1 : 0);
    const int exc_off = poff + pbytes;
    const int isz = exc_index_size(exc);
    if (exc <= kExcBitmap) {
      if (e0) B[exc_off + rank0] = (uint8_t)p0;
      if (e1) B[exc_off + rank1] = (uint8_t)p1;
    } else if (t < 16) {
      const uint32_t n0 = (uint32_t)(b0 >> (4 * t)) & 0xF, n1 = (uint32_t)(b1 >> (4 * t)) & 0xF;
      uint32_t byte = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) byte |= (((n0 >> i) & 1) << (2 * i)) | (((n1 >> i) & 1) << (2 * i + 1));
      B[exc_off + t] = (uint8_t)byte;
    }
    // StreamVByte of the high bits, in position (= rank) order
    const uint32_t h0 = x0 >> bw, h1 = x1 >> bw;
    const int c0 = h0 < (1u << 8) ? 0 : h0 < (1u << 16) ? 1 : h0 < (1u << 24) ? 2 : 3;
    const int c1 = h1 < (1u << 8) ? 0 : h1 < (1u << 16) ? 1 : h1 < (1u << 24) ? 2 : 3;
    const int l0 = e0 ? c0 + 1 : 0, l1 = e1 ? c1 + 1 : 0;
    const uint32_t incl = wave_incl_scan((uint32_t)(l0 + l1));
    const int total = (int)uniform(__shfl(incl, 63, 64));
    const int ctrl = svb_ctrl_count(exc);
    const int svb_off = exc_off + isz;
    const int d0 = svb_off + ctrl + (int)incl - (l0 + l1);
    if (e0) {
      L->codes[rank0] = (uint8_t)c0;
      put_bytes(B + d0, h0, l0);
    }
    if (e1) {
      L->codes[rank1] = (uint8_t)c1;
      put_bytes(B + d0 + l0, h1, l1);
    }
    lds_fence();
    if (t < ctrl) {
      uint32_t cb = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (4 * t + s < exc) cb |= (uint32_t)L->codes[4 * t + s] << (2 * s);
      B[svb_off + t] = (uint8_t)cb;
    }
    const int svb_len = ctrl + total;
    if (t == 0) put_bytes(B + kHdr, (uint32_t)svb_len, 2);
    len = svb_off + svb_len;
  }
  lds_fence();
  return len;
}

__device__ __forceinline__ void copy_out(const EncLds* L, int len, uint8_t* __restrict__ dst) {
  for (int i = lane_id(); i < len; i += 64) dst[i] = L->blk[i];
}

// Encodes block b of a u32 job table into its slot.
__global__ __launch_bounds__(64 * kEncWaves) void k_encode_u32(const uint32_t* __restrict__ values,
                                                              const uint64_t* __restrict__ in_off,
                                                              const uint32_t* __restrict__ in_meta, uint64_t nblocks,
                                                              uint8_t* __restrict__ slots, uint32_t slot_bytes,
                                                              uint32_t* __restrict__ out_meta,
                                                              uint32_t* __restrict__ work_out) {
  __shared__ EncLds lds[kEncWaves];
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint64_t b = (uint64_t)blockIdx.x * kEncWaves + wave;
  if (b >= nblocks) return;
  EncLds* L = &lds[wave];
  const uint32_t meta = in_meta[b];
  const int n = (int)(meta & 0xFF);
  const uint32_t flag = (meta >> 8) & 0xFF;
  const int t = lane_id();
  if (n < 1 || n > kBlock) {  // packBlockScalar: ErrInvalidBuffer (scalar.go:127-129)
    if (t == 0) out_meta[b] = 0xFFFFFFFFu;
    return;
  }
  const uint32_t* v = values + in_off[b];
  uint32_t x0 = (2 * t < n) ? v[2 * t] : 0u;
  uint32_t x1 = (2 * t + 1 < n) ? v[2 * t + 1] : 0u;
  const int len = pack_sub_block(L, 0, flag, x0, x1, n, kTypeU32, false, false, 0);
  copy_out(L, len, slots + b * slot_bytes);
  if (work_out) {
    if (2 * t < n) work_out[in_off[b] + 2 * t] = x0;
    if (2 * t + 1 < n) work_out[in_off[b] + 2 * t + 1] = x1;
  }
  if (t == 0) out_meta[b] = (uint32_t)len | ((uint32_t)n << 24);
}

// uint64.go:302-333 packUint64Scalar
__global__ __launch_bounds__(64 * kEncWaves) void k_encode_u64(const uint64_t* __restrict__ values,
                                                              const uint64_t* __restrict__ in_off,
                                                              const uint32_t* __restrict__ in_meta, uint64_t nblocks,
                                                              uint8_t* __restrict__ slots, uint32_t slot_bytes,
                                                              uint32_t* __restrict__ out_meta) {
  __shared__ EncLds lds[kEncWaves];
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint64_t b = (uint64_t)blockIdx.x * kEncWaves + wave;
  if (b >= nblocks) return;
  EncLds* L = &lds[wave];
  const uint32_t meta = in_meta[b];
  const int n = (int)(meta & 0xFF);
  const uint32_t flag = (meta >> 8) & 0xFF;
  const int t = lane_id();
  if (n < 1 || n > kBlock) {  // packUint64Scalar: ErrInvalidBuffer (uint64.go:304-306)
    if (t == 0) out_meta[b] = 0xFFFFFFFFu;
    return;
  }
  const uint64_t* v = values + in_off[b];
  const bool v0 = 2 * t < n, v1 = 2 * t + 1 < n;
  const uint64_t a0 = v0 ? v[2 * t] : 0ull, a1 = v1 ? v[2 * t + 1] : 0ull;
  const uint64_t mn = wave_reduce_min64(min(v0 ? a0 : ~0ull, v1 ? a1 : ~0ull));
  const uint64_t mx = wave_reduce_max64(max(a0, a1));
  const uint64_t orall = wave_reduce_or64(a0 | a1);
  int len;
  if (!(flag & UTL_NOFOR) && mn > 0 && (mx - mn) < (1ull << 32)) {
    uint32_t x0 = (uint32_t)(a0 - mn), x1 = (uint32_t)(a1 - mn);
    if (!v0) x0 = 0;
    if (!v1) x1 = 0;
    len = pack_sub_block(L, 0, flag | UTL_NOFOR, x0, x1, n, kTypeU64, false, true, mn);
  } else if ((orall >> 32) == 0) {
    uint32_t x0 = (uint32_t)a0, x1 = (uint32_t)a1;
    len = pack_sub_block(L, 0, flag, x0, x1, n, kTypeU64, false, false, 0);
  } else {
    uint32_t l0 = (uint32_t)a0, l1 = (uint32_t)a1, h0 = (uint32_t)(a0 >> 32), h1 = (uint32_t)(a1 >> 32);
    const int len1 = pack_sub_block(L, 0, flag, l0, l1, n, kTypeU64 | kHCombine, true, false, 0);
    const int len2 = pack_sub_block(L, len1, flag, h0, h1, n, kTypeU64, false, false, 0);
    if (t == 0) {
      const uint32_t hd = (uint32_t)L->blk[0] | ((uint32_t)L->blk[1] << 8) | ((uint32_t)L->blk[2] << 16) |
                          ((uint32_t)L->blk[3] << 24);
      const Hdr h = decode_header(hd);
      put_bytes(L->blk + kHdr + (h.exc > 0 ? kSvbLen : 0) + for_base_bytes(h.for_width), (uint32_t)len2, 2);
    }
    lds_fence();
    len = len1 + len2;
  }
  copy_out(L, len, slots + b * slot_bytes);
  if (t == 0) out_meta[b] = (uint32_t)len | ((uint32_t)n << 24);
}

// Little-endian dword of slot bytes [i, i + 4) (slots are 4-byte aligned and
// i + 8 stays inside the slot for every i the compaction asks for).
__device__ __forceinline__ uint32_t slot_u32(const uint8_t* __restrict__ s, uint32_t i) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(s) + (i >> 2);
  return __builtin_amdgcn_alignbyte(w[1], w[0], i & 3);
}

// Concatenates the slots into dst (block b's bytes at dst + boff[b]) with
// aligned dword stores: the output dword d is written by the block holding
// byte 4d; blocks are at least 4 bytes, so its bytes past that block come
// from the next block's slot. One wavefront per block. dst must have room
// for the dword holding the last byte.
__global__ __launch_bounds__(256) void k_compact(const uint8_t* __restrict__ slots, uint32_t slot_bytes,
                                                 const uint64_t* __restrict__ boff, uint64_t nblocks,
                                                 uint8_t* __restrict__ dst) {
  const uint64_t b = (uint64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6);
  if (b >= nblocks) return;
  const uint64_t o = boff[b], o1 = boff[b + 1];
  const uint8_t* s = slots + b * slot_bytes;
  const uint8_t* sn = slots + (b + 1 < nblocks ? b + 1 : b) * slot_bytes;
  uint32_t* d32 = reinterpret_cast<uint32_t*>(dst);
  const uint64_t d0 = (o + 3) >> 2, d1 = (o1 + 3) >> 2;  // dwords whose first byte is in this block
  for (uint64_t d = d0 + lane_id(); d < d1; d += 64) {
    const uint64_t p = 4 * d;            // first byte of the dword, in [o, o1)
    // in <= MaxBlockLength - 1 for every block the encoders produce; the
    // clamp keeps reads inside the slot for an index built over rejected
    // blocks (meta 0xFFFFFFFF), whose bytes are meaningless anyway
    const uint32_t in = (uint32_t)min(p - o, (uint64_t)(slot_bytes - 8));
    uint32_t w = slot_u32(s, in);
    const uint64_t mine = o1 - p;        // bytes of the dword in this block (>= 1)
    if (mine < 4) {
      const uint32_t keep = (1u << (8 * mine)) - 1u;
      w = (w & keep) | (b + 1 < nblocks ? (slot_u32(sn, 0) << (8 * mine)) : 0u);
    }
    d32[d] = w;
  }
}

// Job table for streams cut into 128-value blocks: stream s covers values
// [val_off[s], val_off[s+1]) with flag flags[s]; its blocks are numbered from
// blk_off[s] (blk_off[nstreams] blocks in all, read on the device). A
// grid-stride loop over blocks; each block's stream by binary search (the
// last stream starting at or before it, so empty streams are skipped).
__global__ __launch_bounds__(256) void k_make_jobs(const uint64_t* __restrict__ val_off,
                                                   const uint64_t* __restrict__ blk_off,
                                                   const uint8_t* __restrict__ flags, uint64_t nstreams,
                                                   uint64_t* __restrict__ in_off, uint32_t* __restrict__ in_meta) {
  const uint64_t nblocks = blk_off[nstreams];
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nblocks; k += step) {
    uint64_t lo = 0, hi = nstreams - 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) >> 1;
      if (blk_off[mid] <= k)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint64_t s = lo, i = k - blk_off[s];
    const uint64_t v0 = val_off[s], nv = val_off[s + 1] - v0;
    const uint64_t rest = nv - 128 * i;
    in_off[k] = v0 + 128 * i;
    in_meta[k] = (uint32_t)(rest < 128 ? rest : 128) | ((uint32_t)flags[s] << 8);
  }
}

}  // namespace utl

using namespace utl;

hipError_t utl_launch_encode_u32(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                 uint64_t nblocks, uint8_t* slots, uint32_t slot_bytes, uint32_t* out_meta,
                                 uint32_t* work_out, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t grid = (nblocks + kEncWaves - 1) / kEncWaves;
  k_encode_u32<<<dim3((unsigned)grid), dim3(64 * kEncWaves), 0, stream>>>(values, in_off, in_meta, nblocks, slots,
                                                                            slot_bytes, out_meta, work_out);
  return hipGetLastError();
}

hipError_t utl_launch_encode_u64(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                 uint64_t nblocks, uint8_t* slots, uint32_t slot_bytes, uint32_t* out_meta,
                                 hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t grid = (nblocks + kEncWaves - 1) / kEncWaves;
  k_encode_u64<<<dim3((unsigned)grid), dim3(64 * kEncWaves), 0, stream>>>(values, in_off, in_meta, nblocks, slots,
                                                                            slot_bytes, out_meta);
  return hipGetLastError();
}

hipError_t utl_launch_compact(const uint8_t* slots, uint32_t slot_bytes, const uint64_t* boff, uint64_t nblocks,
                              uint8_t* dst, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  k_compact<<<dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, stream>>>(slots, slot_bytes, boff, nblocks, dst);
  return hipGetLastError();
}

hipError_t utl_launch_make_jobs(const uint64_t* val_off, const uint64_t* blk_off, const uint8_t* flags,
                                uint64_t nstreams, uint64_t* in_off, uint32_t* in_meta, hipStream_t stream) {
  if (nstreams == 0) return hipSuccess;
  k_make_jobs<<<dim3(2048), dim3(256), 0, stream>>>(val_off, blk_off, flags, nstreams, in_off, in_meta);
  return hipGetLastError();
}
