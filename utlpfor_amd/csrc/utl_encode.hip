// utl_encode.hip -- gfx950 encode kernels for UTL-PFOR blocks.
//
// Eight blocks per wavefront: the 8 threads of an aligned 8-lane group own
// one block, and every per-block decision is a reduction over the group
// (DPP), so the decision work of eight blocks runs as vector instructions in
// one wavefront. Blocks are written into fixed-size slots (a scan of the
// block lengths and a compaction pass then produce the concatenated stream)
// or, for uint32 blocks, straight to their place in the stream, with
// offsets found by a decoupled look-back over workgroup tiles.
//
// The steps follow packBlockScalar (scalar.go:126-235) exactly:
//   FOR decision decideFORFull (for.go:95-118) from the group's min/max and
//   the 8 threshold counts of the raw values; per-lane delta with zigzag of
//   all values on any decrease (delta.go:23-47); bit width by the integer
//   cost model over exception counts (select_bitwidth.go:41-121; ties keep
//   the wider width); UTL payload words packed from registers
//   (scalar.go:7-59); exceptions ranked within the group into a sorted
//   position list or a bitmap (exceptions.go:138-186) with their high bits
//   StreamVByte-encoded ("1234" codes, control bytes first).
// uint64 blocks follow packUint64Scalar (uint64.go:302-333): FOR64, fit-in-32
// or two sub-blocks with block2Len patched into the first, each sub-block
// through the same uint32 steps.
#include "utl_device.h"
#include "utl_encode.h"
#include "utl_format.h"

namespace utl {

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- blocks, eight per wavefront ----
//
// Thread j (0..7) of a block's 8-lane group owns UTL lanes 2j and 2j+1, i.e.
// the 16 values at positions 16r + 2j + e, so per-lane delta runs in
// registers and the payload words of its two lanes are packed from
// registers.
//
// Slot layout. uint32: the 32 blocks of a workgroup tile back to back from
// the tile's first slot (offsets from a wave scan of the block lengths and
// the workgroup's wave totals), so the compaction copies each tile's bytes
// in one piece. uint64: block b from byte 0 of slot b. Both are written
// with the writers below, which take any byte address.

constexpr int kEncGrpWaves = 4;
constexpr int kWaveBlocksEnc = 8;  // blocks per wavefront

// Each block's exceptions by rank (their order by position): the high bits
// (32-bit words from byte kExhOff of the block's sec row), and in exm the
// offset of the exception's data bytes among the section's data bytes
// (bits 0-8: at most 127 * 4) and its position (bits 9-15).
// The sec row also stages the block's exception section before it is stored
// (tail_write).
constexpr int kExhOff = 52;   // >= the longest index + control bytes (16 + 32) + 3 alignment bytes + 1
constexpr int kSecRow = 576;  // kExhOff + 4 * 128 high-bit words, rounded up to 16 bytes
struct alignas(16) EncGrpLds {
  uint8_t sec[kWaveBlocksEnc][kSecRow];
  uint16_t exm[kWaveBlocksEnc][kBlock];
};
static_assert(sizeof(EncGrpLds) * kEncGrpWaves <= 32768, "LDS per workgroup");

__device__ __forceinline__ uint32_t step_len(uint32_t x) { return (uint32_t)(bit_len32(x) + 3) >> 2; }

// Per-block counts exc[si] of values whose step index exceeds si, si = 0..7
// (select_bitwidth.go:41-51), as 8-bit fields: si 0..3 in lo, 4..7 in hi.
// Values at invalid positions must be 0. Each thread packs the step indices
// of its 16 values into nibbles; nibble s + (7 - si) has bit 3 set exactly
// when s > si (s <= 8, so no nibble carries).
__device__ __forceinline__ void step_counts(const uint32_t* x0, const uint32_t* x1, uint32_t& lo, uint32_t& hi) {
  uint32_t s_a = 0, s_b = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    s_a |= (step_len(x0[r]) << (8 * r)) | (step_len(x1[r]) << (8 * r + 4));
    s_b |= (step_len(x0[r + 4]) << (8 * r)) | (step_len(x1[r + 4]) << (8 * r + 4));
  }
  uint32_t c_lo = 0, c_hi = 0;
#pragma unroll
  for (int si = 0; si < 8; ++si) {
    const uint32_t add = 0x77777777u - (uint32_t)si * 0x11111111u;
    const uint32_t c = (uint32_t)(__popc((s_a + add) & 0x88888888u) + __popc((s_b + add) & 0x88888888u));
    if (si < 4)
      c_lo |= c << (8 * si);
    else
      c_hi |= c << (8 * (si - 4));
  }
  lo = g8_add(c_lo);  // fields stay <= 128
  hi = g8_add(c_hi);
}

// select_bitwidth.go:56-72 chooseBestFromHist over the counts above, one
// candidate per thread of the group: thread j prices width 4j with exc[j]
// exceptions (for j < max_si; bytes per exception value (k + 1) >> 1 for
// k = max_si - j steps) against the widest step, max_si, which has none.
// The group keeps the lowest cost and, among equal costs, the widest width:
// the reference walks the candidates from the top with a strict '<'.
__device__ __forceinline__ void choose_best_g(uint32_t lo, uint32_t hi, int max_si, int& best_w, int& best_exc) {
  const int j = lane_id() & 7;
  const int e = (int)(((j < 4 ? lo : hi) >> (8 * (j & 3))) & 0xFF);
  const int k = max_si - j;
  const int c = kHdr + (j << 6) + kSvbLen + (e < kExcBitmap ? e : kExcBitmap) + ((e + 3) >> 2) + e * ((k + 1) >> 1);
  // key: cost above 4 bits of 15 - width step (costs stay below 2^12)
  const uint32_t top = ((uint32_t)(kHdr + (max_si << 6)) << 4) | (uint32_t)(15 - max_si);
  const uint32_t mine = ((uint32_t)c << 4) | (uint32_t)(15 - j);
  const uint32_t key = g8_min(j < max_si ? min(mine, top) : top);
  const int si = 15 - (int)(key & 15);
  best_w = si << 2;
  best_exc = si >= max_si ? 0 : (int)(((si < 4 ? lo : hi) >> (8 * (si & 3))) & 0xFF);
}

// ---- the steps of one (sub-)block, per 8-lane group ----
// x0/x1 hold the thread's 16 work values; positions at or past the count
// (ok0/ok1 false) hold 0 on entry and keep 0.

// FOR decision decideFORFull (for.go:95-118), given the min mn and max mx of
// the valid values (n of them): when FOR pays, subtracts the base from the
// valid values and returns the FOR width code (1, 2, 3: a 1, 2 or 4 byte
// base); 0 = no FOR.
// The standard width stdWidth needs the step histogram, except for a full
// block whose values all share one step index s (step_len(mn) ==
// step_len(mx)): every narrower candidate 4j then has all 128 values as
// exceptions and costs 2 + 16 + 32 + 128 * ((k + 1) >> 1) - 64 k >= 50 bytes
// more than width 4s (k = s - j >= 1 steps), so stdWidth = 4s.
__device__ __forceinline__ int enc_for(uint32_t* x0, uint32_t* x1, const bool* ok0, const bool* ok1, uint32_t mn,
                                       uint32_t mx, int n, uint32_t& base) {
  base = 0;
  if (mn == 0) return 0;
  const int for_max = ((bit_len32(mx - mn) + 3) >> 2) << 2;
  const int raw_max = ((bit_len32(mx) + 3) >> 2) << 2;
  if (for_max >= raw_max) return 0;
  int std_w;
  if (n == kBlock && step_len(mn) == step_len(mx)) {
    std_w = raw_max;
  } else {
    uint32_t lo, hi;
    step_counts(x0, x1, lo, hi);
    int unused;
    choose_best_g(lo, hi, (int)step_len(mx), std_w, unused);
  }
  if (for_max >= std_w) return 0;
  base = mn;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    x0[r] = ok0[r] ? x0[r] - base : 0u;
    x1[r] = ok1[r] ? x1[r] - base : 0u;
  }
  const int bytes = (bit_len32(mn) + 7) >> 3;  // >= 1
  return bytes == 1 ? 1 : bytes == 2 ? 2 : 3;
}

// Per-lane delta (delta.go:23-47): row r minus row r-1 in each UTL lane, and
// zigzag of every value when any valid one decreased. Returns the header bits.
__device__ __forceinline__ uint32_t enc_delta(uint32_t* x0, uint32_t* x1, const bool* ok0, const bool* ok1) {
  bool dec = false;
#pragma unroll
  for (int r = kRows - 1; r >= 1; --r) {
    dec |= ok0[r] && x0[r] < x0[r - 1];
    dec |= ok1[r] && x1[r] < x1[r - 1];
    x0[r] = ok0[r] ? x0[r] - x0[r - 1] : 0u;
    x1[r] = ok1[r] ? x1[r] - x1[r - 1] : 0u;
  }
  if (!g8_or(dec ? 1u : 0u)) return kHDelta;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    x0[r] = zz_enc(x0[r]);  // zz_enc(0) = 0 keeps invalid positions at 0
    x1[r] = zz_enc(x1[r]);
  }
  return kHDelta | kHZigZag;
}

// Bit width and exception count (select_bitwidth.go:20-121): the widest step
// with NoPatch, else the cost model.
__device__ __forceinline__ void enc_width(const uint32_t* x0, const uint32_t* x1, bool nopatch, int& bw, int& exc) {
  uint32_t orall = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) orall |= x0[r] | x1[r];
  const int max_si = (int)step_len(g8_or(orall));
  bw = max_si << 2;
  exc = 0;
  if (!nopatch) {
    uint32_t lo, hi;
    step_counts(x0, x1, lo, hi);
    choose_best_g(lo, hi, max_si, bw, exc);
  }
}

// ---- the exception section (exceptions.go:138-186) ----
//
// For a block with exc > 0 exceptions at width bw: the position index (a
// sorted position list up to kExcBitmap exceptions, else the 16-byte
// bitmap), then the StreamVByte control bytes (the code, byte length - 1, of
// rank k in bits 2 (k & 3) of control byte k >> 2), then the data bytes (the
// high bits of each exception, little endian, in position order). A rank is
// the number of exceptions at lower positions. tail_plan ranks the
// exceptions (one pass over the positions, writing each one's high bits and
// data offset and position to LDS by rank, the offsets from scans of
// per-row byte counts over the group) and sizes the section; tail_write assembles
// the section in LDS in rounds of eight ranks, rank 8q + j to thread j, then
// stores it at its byte address as aligned dwords.
struct Tail {
  uint32_t bmw;  // bitmap word j >> 1 (bitmap bytes 2j, 2j+1)
  int isz, ctrl, svb_len;
};

__device__ __forceinline__ uint32_t svb_code(uint32_t h) { return (uint32_t)((bit_len32(h | 1u) - 1) >> 3); }

// Data bytes of the high bits x >> bw of an exception x >= 2^bw (its
// StreamVByte code + 1): (bit_len(x) - bw + 7) >> 3.
__device__ __forceinline__ uint32_t exc_bytes(uint32_t x, int bw) {
  return (uint32_t)(39 - bw - __clz((int)x)) >> 3;
}

__device__ __forceinline__ void tail_plan(EncGrpLds* L, int blk, int j, const uint32_t* x0, const uint32_t* x1,
                                          const bool* ok0, const bool* ok1, int bw, int exc, Tail& t) {
  const uint32_t m = low_mask(bw);
  // bitmap of exception positions: word d holds positions 32d .. 32d + 31
  // (rows 2d, 2d+1; bit 16 (r & 1) + 2j + e), this thread's bits gathered
  // at 16 (r & 1) + e and moved to 2j once
  uint32_t bm[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    bm[r >> 1] |= (ok0[r] && x0[r] > m) ? 1u << (16 * (r & 1)) : 0u;
    bm[r >> 1] |= (ok1[r] && x1[r] > m) ? 2u << (16 * (r & 1)) : 0u;
  }
#pragma unroll
  for (int d = 0; d < 4; ++d) bm[d] = g8_or(bm[d] << (2 * j));
  const uint32_t pre1 = __popc(bm[0]), pre2 = pre1 + __popc(bm[1]), pre3 = pre2 + __popc(bm[2]);
  // Data bytes (code + 1 each) of this thread's exceptions in row r: byte r
  // of lA (rows 0-3) or byte r - 4 of lB, those of its first value in fA,
  // fB. A row holds at most 16 exceptions, so every per-row sum below stays
  // <= 64 and fits its byte.
  uint32_t lA = 0, lB = 0, fA = 0, fB = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const bool e0 = ok0[r] && x0[r] > m, e1 = ok1[r] && x1[r] > m;
    const uint32_t l0 = e0 ? exc_bytes(x0[r], bw) : 0u;
    const uint32_t l = l0 + (e1 ? exc_bytes(x1[r], bw) : 0u);
    if (r < 4) {
      lA |= l << (8 * r);
      fA |= l0 << (8 * r);
    } else {
      lB |= l << (8 * (r - 4));
      fB |= l0 << (8 * (r - 4));
    }
  }
  // per row: the data bytes of the group's threads before this one, and the
  // row's total; then the rows before each row (bytes 0-3 of PA and PB, at
  // most 3 * 64, so no byte carries), sA and sB the totals of rows 0-3, 4-7
  const uint32_t xA = g8_incl_scan(lA) - lA, xB = g8_incl_scan(lB) - lB;
  const uint32_t tA = g8_add(lA), tB = g8_add(lB);
  const uint32_t PA = tA * 0x01010100u, PB = tB * 0x01010100u;
  const uint32_t sA = (PA >> 24) + (tA >> 24), sB = (PB >> 24) + (tB >> 24);
  const uint32_t oA = PA + xA, oB = PB + xB;  // per row: data bytes before this thread's first value (fields <= 255)
  uint32_t* exh = reinterpret_cast<uint32_t*>(L->sec[blk] + kExhOff);
  uint16_t* exm = L->exm[blk];
  const uint32_t below0 = (1u << (2 * j)) - 1u, below1 = (0x10000u << (2 * j)) - 1u;  // bits before 2j, 16 + 2j
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int d = r >> 1;
    const uint32_t pd = d == 0 ? 0u : d == 1 ? pre1 : d == 2 ? pre2 : pre3;
    const uint32_t rank0 = pd + __popc(bm[d] & ((r & 1) ? below1 : below0));
    const bool e0 = ok0[r] && x0[r] > m, e1 = ok1[r] && x1[r] > m;
    const uint32_t D0 = r < 4 ? (oA >> (8 * r)) & 0xFFu : sA + ((oB >> (8 * (r - 4))) & 0xFFu);
    uint32_t v0 = x0[r], v1 = x1[r];
    asm volatile("" : "+v"(v0), "+v"(v1));  // shift again here rather than keep the first pass's results live
    if (e0) {
      exh[rank0] = v0 >> bw;
      exm[rank0] = (uint16_t)(D0 | ((uint32_t)(16 * r + 2 * j) << 9));
    }
    if (e1) {
      const uint32_t l0 = r < 4 ? (fA >> (8 * r)) & 0xFFu : (fB >> (8 * (r - 4))) & 0xFFu;
      const uint32_t rank1 = rank0 + (e0 ? 1u : 0u);
      exh[rank1] = v1 >> bw;
      exm[rank1] = (uint16_t)((D0 + l0) | ((uint32_t)(16 * r + 2 * j + 1) << 9));
    }
  }
  t.isz = exc_index_size(exc);
  t.ctrl = svb_ctrl_count(exc);
  t.svb_len = t.ctrl + (int)(sA + sB);
  t.bmw = (j >> 1) == 0 ? bm[0] : (j >> 1) == 1 ? bm[1] : (j >> 1) == 2 ? bm[2] : bm[3];
  lds_fence();
}

// The section is staged in the block's sec row from byte L0 on, where L0 is
// congruent to T modulo 4, so that the stored dwords are LDS dwords. L0 plus
// the index and control bytes is at most kExhOff: round q writes data bytes
// below byte kExhOff + 4 (8q + 8) (each of the first 8q + 8 ranks has at
// most 4), where the high-bit words of the later rounds begin, and it reads
// its own ranks' words before it writes.
__device__ __forceinline__ void tail_write(EncGrpLds* L, int blk, int j, int exc, const Tail& t,
                                           uint8_t* __restrict__ g, uint64_t T) {
  uint8_t* sec = L->sec[blk];
  const uint32_t* exh = reinterpret_cast<const uint32_t*>(sec + kExhOff);
  const uint16_t* exm = L->exm[blk];
  const int hdr = t.isz + t.ctrl;  // index and control bytes
  const int a = (int)(T & 3);
  const int L0 = (kExhOff - hdr) - (((kExhOff - hdr) - a) & 3);
  const bool list = exc <= kExcBitmap;
  if (!list) {  // bitmap bytes 2j, 2j+1 (position 8q + i is bit i of byte q)
    sec[L0 + 2 * j] = (uint8_t)(t.bmw >> (16 * (j & 1)));
    sec[L0 + 2 * j + 1] = (uint8_t)(t.bmw >> (16 * (j & 1) + 8));
  }
  const int cbase = L0 + t.isz, dbase = L0 + hdr;
  const int rounds = (exc + 7) >> 3;
  uint32_t e = exm[j], h = exh[j];  // rank j (rank 8q + j in round q; every rank read is < 128)
#pragma unroll 1
  for (int q = 0; q < rounds; ++q) {
    const int k = 8 * q + j;
    const int kn = k + 8 < kBlock ? k + 8 : k;
    const uint32_t en = exm[kn], hn = exh[kn];
    const bool valid = k < exc;
    const uint32_t c = valid ? svb_code(h) : 0u;
    const int at = dbase + (int)(e & 0x1FFu);
    if (valid) {
      sec[at] = (uint8_t)h;
      if (c >= 1) sec[at + 1] = (uint8_t)(h >> 8);
      if (c >= 2) sec[at + 2] = (uint8_t)(h >> 16);
      if (c >= 3) sec[at + 3] = (uint8_t)(h >> 24);
      if (list) sec[L0 + k] = (uint8_t)(e >> 9);  // the position list
    }
    // control bytes 2q, 2q + 1: the codes of threads 0..3 and 4..7
    uint32_t cb = c << (2 * (j & 3));
    cb |= g8_dpp<0xB1>(cb);
    cb |= g8_dpp<0x4E>(cb);
    const int ci = 2 * q + (j >> 2);
    if ((j & 3) == 0 && ci < t.ctrl) sec[cbase + ci] = (uint8_t)cb;
    e = en;
    h = hn;
  }
  lds_fence();
  // the section's dwords: whole ones as dword stores, the partial ones at
  // its two ends byte by byte by threads 0 and 1 (their other bytes belong
  // to the payload before it and to the next block)
  const int e_end = a + t.isz + t.svb_len;  // one past the section, from dword 0
  const int i0 = a ? 1 : 0, i1 = e_end >> 2;
  const uint32_t* s32 = reinterpret_cast<const uint32_t*>(sec + (L0 - a));
  uint8_t* g0 = g + (T - (uint64_t)a);
  uint32_t* g32 = reinterpret_cast<uint32_t*>(g0);
#pragma unroll 1
  for (int i = i0 + j; i < i1; i += 8) g32[i] = s32[i];
  // head: bytes [a, min(4, e_end)) of dword 0; tail: bytes [0, e_end & 3) of
  // dword i1 when that is not dword 0 as well
  const bool head = j == 0 && a != 0, tail = j == 1 && (e_end & 3) != 0 && (a == 0 || i1 > 0);
  if (head || tail) {
    const int di = head ? 0 : i1, lo = head ? a : 0, hi = head ? min(4, e_end) : (e_end & 3);
    const uint32_t w = s32[di];
#pragma unroll
    for (int b = 0; b < 3; ++b)
      if (lo + b < hi) g0[4 * di + lo + b] = (uint8_t)(w >> (8 * (lo + b)));
  }
}

// ---- writing a block at an arbitrary byte offset ----
//
// A block of the concatenated stream starts at any byte. Its bytes are
// written in three parts, each byte exactly once and no byte outside the
// block: the header fields (header, svbLen, FOR base: at most 10 bytes) as
// byte stores; the payload as aligned dwords assembled from the lane words
// (each thread holds the words of lanes 2j and 2j+1; the word before a
// thread's first one comes from its neighbour by DPP), with byte stores for
// the partial dwords at its two ends; the exception bytes from LDS the same
// way. Partial dwords at block boundaries are shared with the neighbouring
// blocks, which write their own bytes of them.

__device__ __forceinline__ void put_bytes_g(uint8_t* __restrict__ p, uint32_t v, int n) {
  for (int i = 0; i < n; ++i) p[i] = (uint8_t)(v >> (8 * i));
}

// Payload of lanes 2j, 2j+1 (scalar.go:7-59) written to global byte address
// P (any alignment): word w of lane L is payload dword 16w + L.
template <int BW>
__device__ __forceinline__ void pack_bw_at(const uint32_t* v0, const uint32_t* v1, uint8_t* __restrict__ g,
                                           uint64_t P, int j) {
  if constexpr (BW > 0) {
    constexpr int NW = BW / 4;
    const uint32_t s = (uint32_t)(P & 3);
    uint32_t* a32 = reinterpret_cast<uint32_t*>(g + (P - s));  // aligned dword holding payload byte 0
    uint32_t prev_c = 0;  // this thread's odd word of the previous w
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t a = 0, c = 0;
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int sft = r * BW - 32 * w;
        if (sft > -BW && sft < 32) {
          a |= sft >= 0 ? (v0[r] << sft) : (v0[r] >> (-sft));
          c |= sft >= 0 ? (v1[r] << sft) : (v1[r] >> (-sft));
        }
      }
      const int qe = 16 * w + 2 * j;
      if (s == 0) {
        a32[qe] = a;
        a32[qe + 1] = c;
      } else {
        // payload dword q - 1 for q = qe: lane 2j - 1's word w (the thread
        // before, DPP row_shr:1) or, for j = 0, lane 15's word w - 1 (thread
        // 7 of the group, DPP row_shl:7 of its previous odd word)
        const uint32_t from_left = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)c, 0x111, 0xF, 0xF, true);
        const uint32_t from_7 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)prev_c, 0x107, 0xF, 0xF, true);
        const uint32_t before = j ? from_left : from_7;
        a32[qe + 1] = __builtin_amdgcn_alignbyte(c, a, 4 - s);
        if (qe > 0)
          a32[qe] = __builtin_amdgcn_alignbyte(a, before, 4 - s);
        else
          put_bytes_g(g + P, a, 4 - (int)s);  // payload bytes 0 .. 3 - s
        if (w == NW - 1 && j == 7) put_bytes_g(g + P + 16 * BW - s, c >> (8 * (4 - s)), (int)s);  // the last s
      }
      prev_c = c;
    }
  }
}

__device__ __forceinline__ void pack_any_at(int bw, const uint32_t* v0, const uint32_t* v1, uint8_t* __restrict__ g,
                                            uint64_t P, int j) {
  switch (bw) {
    case 4: pack_bw_at<4>(v0, v1, g, P, j); break;
    case 8: pack_bw_at<8>(v0, v1, g, P, j); break;
    case 12: pack_bw_at<12>(v0, v1, g, P, j); break;
    case 16: pack_bw_at<16>(v0, v1, g, P, j); break;
    case 20: pack_bw_at<20>(v0, v1, g, P, j); break;
    case 24: pack_bw_at<24>(v0, v1, g, P, j); break;
    case 28: pack_bw_at<28>(v0, v1, g, P, j); break;
    case 32: pack_bw_at<32>(v0, v1, g, P, j); break;
    default: break;
  }
}

// ---- decoupled look-back over workgroup tiles (the direct form) ----
//
// lb[0] is the tile ticket counter, lb[1 + t] the status word of tile t
// (value << 2 | flag: 1 the tile's own byte total, 2 the inclusive prefix,
// 3 failed), lb[1 + ntiles] the error word; the host zero-fills all of them
// before the launch. Workgroups take tiles in the order they start (the
// ticket), so every tile a workgroup waits for belongs to a workgroup that is
// already running. Waits are bounded: a tile that does not see its
// predecessors' totals within kSpinLimit polls, or sees a failed one, marks
// itself failed, writes nothing and sets the error word, so the kernel always
// ends.
constexpr uint32_t kLbAgg = 1, kLbPrefix = 2, kLbFailed = 3;
constexpr uint32_t kSpinLimit = 1u << 20;

// A status word carries all its information (flag and value in one 64-bit
// word, nothing else is read because of it), so its loads and stores are
// relaxed device-scope atomics: an acquire would also invalidate the L2 of
// the polling workgroup's XCD on every poll.
__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive byte offset of tile t whose own total is agg, found by the
// whole workgroup (every thread calls both halves). lookback_begin publishes
// the tile's total and issues the loads of the first window; the caller
// does other work (the exception bytes of its blocks) while they are in
// flight; lookback_finish resolves the window. Its 256 threads read the
// status words of the 512 tiles before the window's end at once (two each),
// so one round trip covers 512 tiles (the nearest prefix is almost always in
// the first window; four words each would take the kernel above 128 VGPRs). In a window, the nearest tile
// with a prefix ends the walk and the tiles after it add their totals. A
// window with a tile not yet published is read again (measured: about three
// times per tile when the tile's own wait began right after its totals, as
// older tiles still finish theirs); a window of totals only moves the walk
// back by its width. sh is LDS of 3 * 4 words. ok is cleared on failure.
constexpr int kLbThreads = 64 * 4;
constexpr int kLbPer = 2;                      // status words per thread
constexpr int kLbWindow = kLbThreads * kLbPer;  // tiles per round trip

struct LbWindow {
  uint64_t v[kLbPer];
  int64_t hi;  // the window is tiles [hi - kLbWindow, hi); entry i is tile hi - 1 - i
};

__device__ __forceinline__ void lb_read(const uint64_t* __restrict__ st, LbWindow& w) {
#pragma unroll
  for (int q = 0; q < kLbPer; ++q) {
    const int64_t k = w.hi - 1 - (int64_t)(threadIdx.x + kLbThreads * q);
    w.v[q] = k >= 0 ? lb_load(&st[k]) : (uint64_t)kLbPrefix;  // before tile 0: a prefix of 0
  }
}

__device__ __forceinline__ void lookback_begin(uint64_t* __restrict__ st, uint64_t t, uint64_t agg, LbWindow& w) {
  if (threadIdx.x == 0) lb_store(&st[t], (agg << 2) | (t == 0 ? kLbPrefix : kLbAgg));
  w.hi = (int64_t)t;
  if (t > 0) lb_read(st, w);
}

__device__ uint64_t lookback_finish(uint64_t* __restrict__ st, uint64_t t, uint64_t agg, LbWindow& w, bool& ok,
                                    uint64_t* sh) {
  const int tid = threadIdx.x, wv = tid >> 6;
  uint64_t excl = 0;
  bool good = true;
  uint32_t spins = 0;
  while (w.hi > 0) {
    // the nearest prefix: the lowest entry index holding one
    uint32_t mine = kLbWindow;
#pragma unroll
    for (int q = kLbPer - 1; q >= 0; --q)
      if ((w.v[q] & 3) == kLbPrefix) mine = tid + kLbThreads * q;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mine = min(mine, (uint32_t)__shfl_xor((int)mine, d, 64));
    if ((tid & 63) == 0) sh[wv] = mine;
    __syncthreads();
    uint64_t ip = kLbWindow;
#pragma unroll
    for (int x = 0; x < kLbThreads / 64; ++x) ip = min(ip, sh[x]);
    // the entries after the prefix (indices below ip) must all be totals
    bool missing = false, failed = false;
    uint64_t part = 0;
#pragma unroll
    for (int q = 0; q < kLbPer; ++q) {
      const uint64_t i = (uint64_t)(tid + kLbThreads * q);
      const uint32_t f = (uint32_t)(w.v[q] & 3);
      if (i < ip) {
        missing |= f != kLbAgg;
        failed |= f == kLbFailed;
      }
      if (i <= ip) part += w.v[q] >> 2;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
    const uint64_t mb = __ballot(missing), fb = __ballot(failed);  // (all lanes vote)
    if ((tid & 63) == 0) {
      sh[4 + wv] = part;
      sh[8 + wv] = (mb ? 1u : 0u) | (fb ? 2u : 0u);
    }
    __syncthreads();
    uint64_t sum = 0, bad = 0;
#pragma unroll
    for (int x = 0; x < kLbThreads / 64; ++x) {
      sum += sh[4 + x];
      bad |= sh[8 + x];
    }
    __syncthreads();  // sh is written again in the next round
    if (bad & 2) {
      good = false;
      break;
    }
    if (bad & 1) {  // a tile of the window has not published yet: read it again
      if (++spins > kSpinLimit) {
        good = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      lb_read(st, w);
      continue;
    }
    spins = 0;
    excl += sum;
    if (ip < (uint64_t)kLbWindow) break;  // reached a prefix
    w.hi -= kLbWindow;                    // totals only: the next window back
    if (w.hi > 0) lb_read(st, w);
  }
  ok = good;
  if (tid == 0 && t != 0) lb_store(&st[t], good ? (((excl + agg) << 2) | kLbPrefix) : kLbFailed);
  return excl;
}

// One wavefront's eight blocks of k_encode_u32. FULL: every block of the
// wave has 128 values (no position past a count, the common case of long
// streams), so loads and masks need no predicates.
template <bool DIRECT, bool FULL>
__device__ __forceinline__ void encode_u32_wave(EncGrpLds* L, uint64_t* wg, uint64_t* lbsh, uint32_t wave,
                                                uint64_t tile, uint64_t b, bool mine, bool ok, int n, uint32_t flag,
                                                uint64_t voff, const uint32_t* __restrict__ values, uint64_t nblocks,
                                                uint8_t* __restrict__ dst, uint32_t slot_bytes,
                                                uint32_t* __restrict__ out_meta, uint32_t* __restrict__ work_out,
                                                uint64_t* __restrict__ lb, uint64_t* __restrict__ block_off,
                                                uint64_t dst_cap) {
  const int blk = lane_id() >> 3, j = lane_id() & 7;
  const uint32_t* v = values + voff;
  uint32_t x0[kRows], x1[kRows];
  bool ok0[kRows], ok1[kRows];
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int p = 16 * r + 2 * j;
    ok0[r] = FULL || p < n;
    ok1[r] = FULL || p + 1 < n;
    x0[r] = ok0[r] ? v[p] : 0u;
    x1[r] = ok1[r] ? v[p + 1] : 0u;
    if (ok0[r]) {
      mn = min(mn, x0[r]);
      mx = max(mx, x0[r]);
    }
    if (ok1[r]) {
      mn = min(mn, x1[r]);
      mx = max(mx, x1[r]);
    }
  }
  uint32_t hflags = kTypeU32 | ((flag & UTL_SPECIAL) << 14);
  uint32_t base = 0;
  const int fw = (flag & UTL_NOFOR) ? 0 : enc_for(x0, x1, ok0, ok1, g8_min(mn), g8_max(mx), n, base);
  hflags |= (uint32_t)fw << 15;
  if (flag & UTL_DELTA) hflags |= enc_delta(x0, x1, ok0, ok1);
  int bw, exc;
  enc_width(x0, x1, (flag & UTL_NOPATCH) != 0, bw, exc);
  if (work_out && ok) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      if (ok0[r]) work_out[voff + 16 * r + 2 * j] = x0[r];
      if (ok1[r]) work_out[voff + 16 * r + 2 * j + 1] = x1[r];
    }
  }

  const bool has_exc = exc > 0;
  const int fbb = for_base_bytes(fw);
  const int poff = kHdr + (has_exc ? kSvbLen : 0) + fbb;
  const uint32_t m = low_mask(bw);

  // ---- the block's length and its offset in the workgroup's tile: the
  // wave's blocks in order, the waves in order. The direct form finds the
  // length before the exception bytes are assembled (their StreamVByte data
  // bytes are a sum of code lengths), publishes the tile's total and reads
  // the look-back's first window while that work runs ----
  const int isz = exc_index_size(exc);
  int len = 0;
  uint32_t blen = 0, incl = 0;
  uint64_t agg = 0, before = 0;
  LbWindow lbw;
  auto tile_scan = [&]() {
    blen = ok ? (uint32_t)len : 0u;
    incl = wave_group_scan(blen);
    if (lane_id() == 63) wg[1 + wave] = incl;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kEncGrpWaves; ++w) {
      agg += wg[1 + w];
      before += w < (int)wave ? wg[1 + w] : 0;
    }
  };
  if (DIRECT) {
    uint32_t eb = 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      eb += (ok0[r] && x0[r] > m) ? exc_bytes(x0[r], bw) : 0u;
      eb += (ok1[r] && x1[r] > m) ? exc_bytes(x1[r], bw) : 0u;
    }
    len = poff + 16 * bw + (has_exc ? isz + svb_ctrl_count(exc) + (int)g8_add(eb) : 0);
    tile_scan();
    lookback_begin(lb + 1, tile, agg, lbw);
  }

  // ---- exceptions (exceptions.go:138-186) ----
  Tail tail;
  if (has_exc) tail_plan(L, blk, j, x0, x1, ok0, ok1, bw, exc, tail);
  const int svb_len = has_exc ? tail.svb_len : 0;
  if (!DIRECT) {
    len = poff + 16 * bw + (has_exc ? isz + svb_len : 0);
    tile_scan();
  }

  uint64_t tbase = tile * (uint64_t)(kEncGrpWaves * kWaveBlocksEnc) * slot_bytes;  // slot form: the tile's slots
  if (DIRECT) {
    bool good;
    const uint64_t lbase = lookback_finish(lb + 1, tile, agg, lbw, good, lbsh);
    if (!good && threadIdx.x == 0) atomicOr((unsigned long long*)(lb + 1 + (nblocks + 31) / 32), 1ull);
    if (!good) return;
    tbase = lbase;
  }
  if (!mine) return;
  const uint64_t o = tbase + before + (incl - blen);
  // a block that would end past dst_cap writes nothing (the caller sizes
  // dst for the longest blocks, so only a wrong offset could get here)
  const bool fits = !DIRECT || (o <= dst_cap && blen <= dst_cap - o);
  if (!fits && j == 0) atomicOr((unsigned long long*)(lb + 1 + (nblocks + 31) / 32), 2ull);
  if (ok && fits) {
    // header fields (bytes [0, poff)): header, svbLen, FOR base
    const uint32_t hw = encode_header(n, bw, exc, hflags);
    const uint32_t w1 = has_exc ? ((uint32_t)svb_len | (base << 16)) : base;
    const uint32_t w2 = has_exc ? (fbb > 2 ? base >> 16 : 0u) : 0u;
    const uint32_t hsel = j < 4 ? hw : w1;
    if (j < poff) dst[o + j] = (uint8_t)(hsel >> (8 * (j & 3)));
    if (j + 8 < poff) dst[o + j + 8] = (uint8_t)(w2 >> (8 * j));
    uint32_t p0[kRows], p1[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      p0[r] = x0[r] & m;
      p1[r] = x1[r] & m;
    }
    pack_any_at(bw, p0, p1, dst, o + poff, j);
    if (has_exc) tail_write(L, blk, j, exc, tail, dst, o + poff + 16 * bw);
  }
  if (j == 0) {
    out_meta[b] = ok ? ((uint32_t)len | ((uint32_t)n << 24)) : 0xFFFFFFFFu;
    if (DIRECT) {
      block_off[b] = o;
      if (b + 1 == nblocks) block_off[nblocks] = o + blen;
    }
  }
}

// Encodes the blocks of a u32 job table, eight per wavefront, following
// packBlockScalar (scalar.go:126-235): FOR decision (for.go:95-118), per-lane
// delta with zigzag of all values on any decrease (delta.go:23-47), bit
// width by the cost model (select_bitwidth.go:41-121), UTL payload,
// exceptions as a sorted position list or a bitmap with their high bits
// StreamVByte-encoded (exceptions.go:138-186).
// Slot form (!DIRECT): block b goes to slot b (slot_bytes apart) of dst.
// Direct form: the workgroup's 32 blocks go straight to their place in the
// concatenated stream in dst, found by the look-back over tiles; block_off[b]
// receives each block's offset (block_off[nblocks] the stream's length).
template <bool DIRECT>
__global__ __launch_bounds__(64 * kEncGrpWaves, 4) void k_encode_u32(const uint32_t* __restrict__ values,
                                                                  const uint64_t* __restrict__ in_off,
                                                                  const uint32_t* __restrict__ in_meta,
                                                                  uint64_t nblocks, uint8_t* __restrict__ dst,
                                                                  uint32_t slot_bytes, uint32_t* __restrict__ out_meta,
                                                                  uint32_t* __restrict__ work_out,
                                                                  uint64_t* __restrict__ lb,
                                                                  uint64_t* __restrict__ block_off,
                                                                  uint64_t dst_cap) {
  __shared__ EncGrpLds lds[kEncGrpWaves];
  __shared__ uint64_t wg[1 + kEncGrpWaves];  // tile ticket, wave totals
  __shared__ uint64_t lbsh[12];               // the look-back's reductions
  static_assert(64 * kEncGrpWaves == kLbThreads, "the look-back runs on the whole workgroup");
  const uint32_t wave = uniform(threadIdx.x >> 6);
  uint64_t tile = blockIdx.x;
  if (DIRECT) {
    if (threadIdx.x == 0) wg[0] = atomicAdd((unsigned long long*)lb, 1ull);
    __syncthreads();
    tile = wg[0];
  }
  const uint64_t b0 = (tile * kEncGrpWaves + wave) * kWaveBlocksEnc;  // (every wave reaches the barrier)
  EncGrpLds* L = &lds[wave];
  const bool mine = b0 + (lane_id() >> 3) < nblocks;
  // (a lane without a block reads block b0's job, or the last block's when
  // the whole wave is past the end: every read stays inside the tables)
  const uint64_t b = mine ? b0 + (lane_id() >> 3) : min(b0, nblocks - 1);
  const uint32_t meta = in_meta[b];
  const uint32_t flag = (meta >> 8) & 0xFF;
  const int n_req = (int)(meta & 0xFF);
  const bool ok = mine && n_req >= 1 && n_req <= kBlock;  // else ErrInvalidBuffer (scalar.go:127-129)
  const int n = ok ? n_req : 0;
  const uint64_t voff = in_off[b];
  if (__ballot(!ok || n != kBlock) == 0)
    encode_u32_wave<DIRECT, true>(L, wg, lbsh, wave, tile, b, mine, ok, n, flag, voff, values, nblocks, dst,
                                  slot_bytes, out_meta, work_out, lb, block_off, dst_cap);
  else
    encode_u32_wave<DIRECT, false>(L, wg, lbsh, wave, tile, b, mine, ok, n, flag, voff, values, nblocks, dst,
                                   slot_bytes, out_meta, work_out, lb, block_off, dst_cap);
}

// ---- the sized form of the uint32 encode: a size pass, the index scan, a write pass ----
//
// The size pass makes every decision of packBlockScalar for a block (FOR,
// zigzag, bit width, exception count) and its length, from which the
// StreamVByte data bytes follow (a byte per exception plus its code). It
// writes the meta word and an 8-byte plan: bw | FOR width code << 8 |
// zigzag << 10 | exception count << 16, then the FOR base. The caller scans
// the meta words (the block offsets); the write pass reloads the values,
// applies the planned FOR and delta, and packs the block at its offset.
// Nothing is staged or copied, and neither pass waits on another workgroup.

__device__ __forceinline__ void enc_load(const uint32_t* __restrict__ v, int n, int j, bool full, uint32_t* x0,
                                         uint32_t* x1, bool* ok0, bool* ok1) {
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int p = 16 * r + 2 * j;
    ok0[r] = full || p < n;
    ok1[r] = full || p + 1 < n;
    x0[r] = ok0[r] ? v[p] : 0u;
    x1[r] = ok1[r] ? v[p + 1] : 0u;
  }
}

template <bool FULL>
__device__ __forceinline__ void size_u32_wave(uint64_t b, bool mine, bool ok, int n, uint32_t flag, uint64_t voff,
                                              const uint32_t* __restrict__ values, uint32_t* __restrict__ out_meta,
                                              uint2* __restrict__ plan) {
  const int j = lane_id() & 7;
  uint32_t x0[kRows], x1[kRows];
  bool ok0[kRows], ok1[kRows];
  enc_load(values + voff, n, j, FULL, x0, x1, ok0, ok1);
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (ok0[r]) {
      mn = min(mn, x0[r]);
      mx = max(mx, x0[r]);
    }
    if (ok1[r]) {
      mn = min(mn, x1[r]);
      mx = max(mx, x1[r]);
    }
  }
  uint32_t base = 0;
  const int fw = (flag & UTL_NOFOR) ? 0 : enc_for(x0, x1, ok0, ok1, g8_min(mn), g8_max(mx), n, base);
  const uint32_t zz = (flag & UTL_DELTA) ? ((enc_delta(x0, x1, ok0, ok1) & kHZigZag) ? 1u : 0u) : 0u;
  int bw, exc;
  enc_width(x0, x1, (flag & UTL_NOPATCH) != 0, bw, exc);
  const uint32_t m = low_mask(bw);
  uint32_t eb = 0;  // StreamVByte data bytes of this thread's exceptions
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    eb += (ok0[r] && x0[r] > m) ? exc_bytes(x0[r], bw) : 0u;
    eb += (ok1[r] && x1[r] > m) ? exc_bytes(x1[r], bw) : 0u;
  }
  eb = g8_add(eb);
  const bool has_exc = exc > 0;
  const int poff = kHdr + (has_exc ? kSvbLen : 0) + for_base_bytes(fw);
  const int len = poff + 16 * bw + (has_exc ? exc_index_size(exc) + svb_ctrl_count(exc) + (int)eb : 0);
  if (j == 0 && mine) {
    out_meta[b] = ok ? ((uint32_t)len | ((uint32_t)n << 24)) : 0xFFFFFFFFu;
    plan[b] = make_uint2((uint32_t)bw | ((uint32_t)fw << 8) | (zz << 10) | ((uint32_t)exc << 16), base);
  }
}

// A block that would end past dst_cap (only an index that is not the scan
// of the size pass's meta words can place it there) is not written.
template <bool FULL>
__device__ __forceinline__ void write_u32_wave(EncGrpLds* L, bool ok, int n, uint32_t flag, uint64_t voff,
                                               const uint32_t* __restrict__ values, uint2 pl, uint64_t o,
                                               uint8_t* __restrict__ dst, uint64_t dst_cap) {
  const int blk = lane_id() >> 3, j = lane_id() & 7;
  uint32_t x0[kRows], x1[kRows];
  bool ok0[kRows], ok1[kRows];
  enc_load(values + voff, n, j, FULL, x0, x1, ok0, ok1);
  const int bw = (int)(pl.x & 0xFF), fw = (int)((pl.x >> 8) & 3), exc = (int)(pl.x >> 16);
  const uint32_t base = pl.y;
  uint32_t hflags = kTypeU32 | ((flag & UTL_SPECIAL) << 14) | ((uint32_t)fw << 15);
  if (fw) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      x0[r] = ok0[r] ? x0[r] - base : 0u;
      x1[r] = ok1[r] ? x1[r] - base : 0u;
    }
  }
  if (flag & UTL_DELTA) {
    // per-lane delta (delta.go:23-47), zigzag of every value as planned
#pragma unroll
    for (int r = kRows - 1; r >= 1; --r) {
      x0[r] = ok0[r] ? x0[r] - x0[r - 1] : 0u;
      x1[r] = ok1[r] ? x1[r] - x1[r - 1] : 0u;
    }
    hflags |= kHDelta;
    if ((pl.x >> 10) & 1) {
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        x0[r] = zz_enc(x0[r]);
        x1[r] = zz_enc(x1[r]);
      }
      hflags |= kHZigZag;
    }
  }
  const bool has_exc = exc > 0;
  Tail tail;
  if (has_exc) tail_plan(L, blk, j, x0, x1, ok0, ok1, bw, exc, tail);
  const int fbb = for_base_bytes(fw);
  const int poff = kHdr + (has_exc ? kSvbLen : 0) + fbb;
  const int svb_len = has_exc ? tail.svb_len : 0;
  const uint64_t len = (uint64_t)(poff + 16 * bw + (has_exc ? exc_index_size(exc) + svb_len : 0));
  if (!ok || o > dst_cap || len > dst_cap - o) return;
  // header fields (bytes [0, poff)): header, svbLen, FOR base
  const uint32_t hw = encode_header(n, bw, exc, hflags);
  const uint32_t w1 = has_exc ? ((uint32_t)svb_len | (base << 16)) : base;
  const uint32_t w2 = has_exc ? (fbb > 2 ? base >> 16 : 0u) : 0u;
  const uint32_t hsel = j < 4 ? hw : w1;
  if (j < poff) dst[o + j] = (uint8_t)(hsel >> (8 * (j & 3)));
  if (j + 8 < poff) dst[o + j + 8] = (uint8_t)(w2 >> (8 * j));
  const uint32_t m = low_mask(bw);
  uint32_t p0[kRows], p1[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    p0[r] = x0[r] & m;
    p1[r] = x1[r] & m;
  }
  pack_any_at(bw, p0, p1, dst, o + poff, j);
  if (has_exc) tail_write(L, blk, j, exc, tail, dst, o + poff + 16 * bw);
}

// The job of lane group (lane >> 3) of wave `wave` of the workgroup: block
// b (a lane without a block reads block b0's job, or the last block's when
// the whole wave is past the end, so every read stays inside the tables).
struct EncJob {
  uint64_t b, voff;
  uint32_t flag;
  int n;
  bool mine, ok, full;
};

__device__ __forceinline__ EncJob enc_job(const uint64_t* __restrict__ in_off, const uint32_t* __restrict__ in_meta,
                                          uint64_t nblocks, uint64_t b0) {
  EncJob e;
  e.mine = b0 + (lane_id() >> 3) < nblocks;
  e.b = e.mine ? b0 + (lane_id() >> 3) : min(b0, nblocks - 1);
  const uint32_t meta = in_meta[e.b];
  e.flag = (meta >> 8) & 0xFF;
  const int n_req = (int)(meta & 0xFF);
  e.ok = e.mine && n_req >= 1 && n_req <= kBlock;  // else ErrInvalidBuffer (scalar.go:127-129)
  e.n = e.ok ? n_req : 0;
  e.voff = in_off[e.b];
  e.full = __ballot(!e.ok || e.n != kBlock) == 0;
  return e;
}

__global__ __launch_bounds__(256) void k_encode_u32_size(const uint32_t* __restrict__ values,
                                                         const uint64_t* __restrict__ in_off,
                                                         const uint32_t* __restrict__ in_meta, uint64_t nblocks,
                                                         uint32_t* __restrict__ out_meta, uint2* __restrict__ plan) {
  const uint64_t b0 = ((uint64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6)) * kWaveBlocksEnc;
  if (b0 >= nblocks) return;  // (the whole wave)
  const EncJob e = enc_job(in_off, in_meta, nblocks, b0);
  if (e.full)
    size_u32_wave<true>(e.b, e.mine, e.ok, e.n, e.flag, e.voff, values, out_meta, plan);
  else
    size_u32_wave<false>(e.b, e.mine, e.ok, e.n, e.flag, e.voff, values, out_meta, plan);
}

__global__ __launch_bounds__(64 * kEncGrpWaves) void k_encode_u32_write(
    const uint32_t* __restrict__ values, const uint64_t* __restrict__ in_off, const uint32_t* __restrict__ in_meta,
    uint64_t nblocks, const uint2* __restrict__ plan, const uint64_t* __restrict__ block_off, uint8_t* __restrict__ dst,
    uint64_t dst_cap) {
  __shared__ EncGrpLds lds[kEncGrpWaves];
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint64_t b0 = ((uint64_t)blockIdx.x * kEncGrpWaves + wave) * kWaveBlocksEnc;
  if (b0 >= nblocks) return;  // (the whole wave; no workgroup barrier follows)
  const EncJob e = enc_job(in_off, in_meta, nblocks, b0);
  const uint2 pl = plan[e.b];
  const uint64_t o = block_off[e.b];
  if (e.full)
    write_u32_wave<true>(&lds[wave], e.ok, e.n, e.flag, e.voff, values, pl, o, dst, dst_cap);
  else
    write_u32_wave<false>(&lds[wave], e.ok, e.n, e.flag, e.voff, values, pl, o, dst, dst_cap);
}

// ---- uint64 blocks, eight per wavefront ----

template <int Ctrl>
__device__ __forceinline__ uint64_t g8_dpp64(uint64_t x) {
  return ((uint64_t)g8_dpp<Ctrl>((uint32_t)(x >> 32)) << 32) | g8_dpp<Ctrl>((uint32_t)x);
}
__device__ __forceinline__ uint64_t g8_min64(uint64_t x) {
  x = min(x, g8_dpp64<0xB1>(x));
  x = min(x, g8_dpp64<0x4E>(x));
  return min(x, g8_dpp64<0x141>(x));
}
__device__ __forceinline__ uint64_t g8_max64(uint64_t x) {
  x = max(x, g8_dpp64<0xB1>(x));
  x = max(x, g8_dpp64<0x4E>(x));
  return max(x, g8_dpp64<0x141>(x));
}
__device__ __forceinline__ uint64_t g8_or64(uint64_t x) {
  return ((uint64_t)g8_or((uint32_t)(x >> 32)) << 32) | g8_or((uint32_t)x);
}

// A sub-block's fields before its payload: header, svbLen (with
// exceptions), FOR base (fbb bytes: 0, 1, 2, 4, or 8 for FOR64), block2Len
// (the first sub-block of a combine block).
struct SubHead {
  uint32_t hw, svb_len, b2len;
  uint64_t base;
  int fbb, poff;
  bool has_exc, wrote;
};

__device__ __forceinline__ uint32_t head_byte(const SubHead& h, int k) {
  if (k < 4) return (h.hw >> (8 * k)) & 0xFFu;
  k -= 4;
  if (h.has_exc) {
    if (k < 2) return (h.svb_len >> (8 * k)) & 0xFFu;
    k -= 2;
  }
  if (k < h.fbb) return (uint32_t)(h.base >> (8 * k)) & 0xFFu;
  return (h.b2len >> (8 * (k - h.fbb))) & 0xFFu;
}

// Bytes [0, poff) of a sub-block at byte o of g (poff <= 14): two per thread.
__device__ __forceinline__ void put_head_at(const SubHead& h, uint8_t* __restrict__ g, uint64_t o, int j) {
  if (j < h.poff) g[o + j] = (uint8_t)head_byte(h, j);
  if (j + 8 < h.poff) g[o + j + 8] = (uint8_t)head_byte(h, j + 8);
}

// One 32-bit sub-block of a uint64 block (packBlockScalar's steps, as for a
// uint32 block) with header type bits tflags: its payload and exception
// section written at byte o of g when `write` and the sub-block ends at or
// before byte `limit` (h.wrote); its header fields returned (written by the
// caller, which knows block2Len last). for64: the values are already minus
// the FOR64 base `base64` and the sub-block carries it.
__device__ __forceinline__ SubHead enc_sub_u64(EncGrpLds* L, int blk, int j, uint32_t* x0, uint32_t* x1,
                                               const bool* ok0, const bool* ok1, int n, uint32_t flag,
                                               uint32_t tflags, bool for64, uint64_t base64, bool combine,
                                               uint8_t* __restrict__ g, uint64_t o, bool write,
                                               uint64_t limit) {
  uint32_t hflags = tflags | ((flag & UTL_SPECIAL) << 14);
  int fw = 0;
  uint32_t base = 0;
  if (!for64 && !(flag & UTL_NOFOR)) {
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      if (ok0[r]) {
        mn = min(mn, x0[r]);
        mx = max(mx, x0[r]);
      }
      if (ok1[r]) {
        mn = min(mn, x1[r]);
        mx = max(mx, x1[r]);
      }
    }
    fw = enc_for(x0, x1, ok0, ok1, g8_min(mn), g8_max(mx), n, base);
  }
  hflags |= (for64 ? 3u : (uint32_t)fw) << 15;
  if (flag & UTL_DELTA) hflags |= enc_delta(x0, x1, ok0, ok1);
  int bw, exc;
  enc_width(x0, x1, (flag & UTL_NOPATCH) != 0, bw, exc);
  SubHead h;
  h.has_exc = exc > 0;
  h.hw = encode_header(n, bw, exc, hflags);
  h.base = for64 ? base64 : base;
  h.fbb = for64 ? kFor64 : for_base_bytes(fw);
  h.b2len = 0;
  h.poff = kHdr + (h.has_exc ? kSvbLen : 0) + h.fbb + (combine ? kB2Len : 0);
  Tail tail;
  if (h.has_exc) tail_plan(L, blk, j, x0, x1, ok0, ok1, bw, exc, tail);
  h.svb_len = h.has_exc ? (uint32_t)tail.svb_len : 0u;
  const uint64_t len = (uint64_t)(h.poff + 16 * bw + (h.has_exc ? exc_index_size(exc) + (int)h.svb_len : 0));
  h.wrote = write && o <= limit && len <= limit - o;
  if (h.wrote) {
    const uint32_t m = low_mask(bw);
    uint32_t p0[kRows], p1[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      p0[r] = x0[r] & m;
      p1[r] = x1[r] & m;
    }
    pack_any_at(bw, p0, p1, g, o + h.poff, j);
    if (h.has_exc) tail_write(L, blk, j, exc, tail, g, o + h.poff + 16 * bw);
  }
  return h;
}

__device__ __forceinline__ int sub_len(const SubHead& h) {
  const int bw = (int)((h.hw >> 8) & 0x1F) << 2, exc = (int)(h.hw >> 24);
  return h.poff + 16 * bw + (h.has_exc ? exc_index_size(exc) + (int)h.svb_len : 0);
}

// Encodes the blocks of a u64 job table, eight per wavefront, following
// packUint64Scalar (uint64.go:302-333): FOR64 when the minimum is above 0
// and max - min < 2^32 (one sub-block of value - min, flag NoFOR, the 8-byte
// base); fit-32 when every value is below 2^32 (one sub-block); else
// combine: the low halves, then the high halves, each a sub-block with the
// block's flag, the first carrying the second's length. MODE 0: block b
// from byte 0 of slot b of dst; MODE 1: lengths only (out_meta); MODE 2:
// block b at block_off[b] of dst (dst_cap bytes; the scan of MODE 1's meta
// words), a sub-block that would end past dst_cap not written.
constexpr int kEncSlots = 0, kEncSize = 1, kEncWrite = 2;

template <int MODE>
__global__ __launch_bounds__(64 * kEncGrpWaves) void k_encode_u64(const uint64_t* __restrict__ values,
                                                                  const uint64_t* __restrict__ in_off,
                                                                  const uint32_t* __restrict__ in_meta,
                                                                  uint64_t nblocks, uint8_t* __restrict__ dst,
                                                                  uint32_t slot_bytes,
                                                                  const uint64_t* __restrict__ block_off,
                                                                  uint64_t dst_cap, uint32_t* __restrict__ out_meta) {
  __shared__ EncGrpLds lds[kEncGrpWaves];
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint64_t b0 = ((uint64_t)blockIdx.x * kEncGrpWaves + wave) * kWaveBlocksEnc;
  if (b0 >= nblocks) return;
  EncGrpLds* L = &lds[wave];
  const int blk = lane_id() >> 3, j = lane_id() & 7;
  const bool mine = b0 + blk < nblocks;
  const uint64_t b = mine ? b0 + blk : b0;  // a lane without a block reads block b0's job
  const uint32_t meta = in_meta[b];
  const uint32_t flag = (meta >> 8) & 0xFF;
  const int n_req = (int)(meta & 0xFF);
  const bool ok = mine && n_req >= 1 && n_req <= kBlock;  // else ErrInvalidBuffer (uint64.go:304-306)
  const int n = ok ? n_req : 0;
  const uint64_t* v = values + in_off[b];

  uint64_t a0[kRows], a1[kRows];
  bool ok0[kRows], ok1[kRows];
  uint64_t mn = ~0ull, mx = 0, orv = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int p = 16 * r + 2 * j;
    ok0[r] = p < n;
    ok1[r] = p + 1 < n;
    a0[r] = ok0[r] ? v[p] : 0ull;
    a1[r] = ok1[r] ? v[p + 1] : 0ull;
    if (ok0[r]) mn = min(mn, a0[r]);
    if (ok1[r]) mn = min(mn, a1[r]);
    mx = max(mx, max(a0[r], a1[r]));
    orv |= a0[r] | a1[r];
  }
  mn = g8_min64(mn);
  mx = g8_max64(mx);
  orv = g8_or64(orv);
  const bool for64 = !(flag & UTL_NOFOR) && mn > 0 && mx - mn < (1ull << 32);
  const bool combine = !for64 && (orv >> 32) != 0;
  uint32_t x0[kRows], x1[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    x0[r] = ok0[r] ? (uint32_t)(for64 ? a0[r] - mn : a0[r]) : 0u;
    x1[r] = ok1[r] ? (uint32_t)(for64 ? a1[r] - mn : a1[r]) : 0u;
  }
  const uint64_t o = MODE == kEncSlots ? b * slot_bytes : MODE == kEncWrite ? block_off[b] : 0;
  const bool wr = MODE != kEncSize;
  const uint64_t limit = MODE == kEncWrite ? dst_cap : ~0ull;
  SubHead h1 = enc_sub_u64(L, blk, j, x0, x1, ok0, ok1, n, for64 ? flag | UTL_NOFOR : flag,
                           kTypeU64 | (combine ? kHCombine : 0u), for64, mn, combine, dst, o, wr && ok, limit);
  int len = sub_len(h1);
  if (combine && ok) {  // the groups of combine blocks: the high halves after the first sub-block
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      x0[r] = (uint32_t)(a0[r] >> 32);
      x1[r] = (uint32_t)(a1[r] >> 32);
    }
    lds_fence();  // the first sub-block's exception bytes are written out
    const SubHead h2 = enc_sub_u64(L, blk, j, x0, x1, ok0, ok1, n, flag, kTypeU64, false, 0, false, dst, o + len,
                                   wr, limit);
    if (h2.wrote) put_head_at(h2, dst, o + len, j);
    h1.b2len = (uint32_t)sub_len(h2);
    len += (int)h1.b2len;
  }
  if (h1.wrote) put_head_at(h1, dst, o, j);
  if (MODE != kEncWrite && j == 0 && mine) out_meta[b] = ok ? ((uint32_t)len | ((uint32_t)n << 24)) : 0xFFFFFFFFu;
}

// Concatenates slots of one block each (the uint64 layout) into dst (block
// b's bytes at dst + boff[b]). Each block writes the aligned dwords inside
// its byte range with dword stores and its bytes of the partial dwords at
// its two ends with byte stores (the neighbouring blocks write the other
// bytes of those), so it reads only its own slot bytes and nothing past the
// stream is written. Eight lanes per block, eight blocks per wavefront: the
// index entries and the slot's first dword (the block starts after any
// zero bytes an encoder put before it; its first byte, the count, is never
// 0) are loaded together, then every slot dword the block needs (up to 64
// per round, eight per lane, each as two dword loads aligned in registers).
constexpr int kCompactPer = 8;  // output dwords per lane per round

__global__ __launch_bounds__(256) void k_compact(const uint8_t* __restrict__ slots, uint32_t slot_bytes,
                                                 const uint64_t* __restrict__ boff, uint64_t nblocks,
                                                 uint8_t* __restrict__ dst) {
  const uint64_t b = ((uint64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6)) * 8 + (lane_id() >> 3);
  if (b >= nblocks) return;
  const int l8 = lane_id() & 7;
  const uint8_t* s = slots + b * slot_bytes;
  const uint32_t head = *reinterpret_cast<const uint32_t*>(s);
  const uint64_t o = boff[b], o1 = boff[b + 1];
  // where the block starts in its slot: after the encoders' zero bytes
  // (0..3); a block's first byte, its count, is never 0
  const uint32_t sh = head ? min((uint32_t)__builtin_ctz(head) >> 3, 3u) : 3u;
  // whole dwords [d0, d1) of dst inside [o, o1) (blocks are at least 4 bytes)
  const uint64_t d0 = (o + 3) >> 2, d1 = o1 >> 2;
  const int ndw = d1 > d0 ? (int)(d1 - d0) : 0;
  uint32_t* d32 = reinterpret_cast<uint32_t*>(dst);
  const uint32_t* s32 = reinterpret_cast<const uint32_t*>(s);
  // output dword d0 + k is slot bytes [in0 + 4k, in0 + 4k + 4): slot dwords
  // (in0 >> 2) + k and the one after, shifted by in0 & 3
  const uint32_t in0 = (uint32_t)(4 * d0 - o + sh);
  const uint32_t q0 = in0 >> 2, r = in0 & 3;
  // sh + (p - o) <= 3 + MaxBlockLength - 1 for every block the encoders
  // produce; the clamp keeps reads inside the slot for an index built over
  // rejected blocks (meta 0xFFFFFFFF), whose bytes are meaningless anyway
  const uint32_t qmax = slot_bytes / 4 - 1;
  for (int k0 = 0; k0 < ndw; k0 += 8 * kCompactPer) {
    uint32_t lo[kCompactPer], hi[kCompactPer];
#pragma unroll
    for (int u = 0; u < kCompactPer; ++u) {
      const int k = k0 + 8 * u + l8;
      lo[u] = hi[u] = 0;
      if (k < ndw) {
        const uint32_t q = min(q0 + (uint32_t)k, qmax);
        lo[u] = s32[q];
        hi[u] = s32[min(q + 1, qmax)];
      }
    }
#pragma unroll
    for (int u = 0; u < kCompactPer; ++u) {
      const int k = k0 + 8 * u + l8;
      if (k < ndw) d32[d0 + k] = __builtin_amdgcn_alignbyte(hi[u], lo[u], r);
    }
  }
  // the bytes before d0 and from d1 on: lanes 0..2 the first, 3..5 the last
  const uint64_t p = l8 < 3 ? o + l8 : 4 * d1 + (uint64_t)(l8 - 3);
  const uint64_t lim = l8 < 3 ? 4 * d0 : o1;
  if (l8 < 6 && p < lim) dst[p] = s[min((uint32_t)(sh + (p - o)), slot_bytes - 1)];
}

// The slot form of the uint32 encode packs the blocks of each workgroup tile
// back to back from the tile's first slot, so concatenating the stream is
// one shifted copy per tile: tile t's bytes [0, boff[b1] - boff[b0]) to dst +
// boff[b0]. The aligned dwords inside that range are written whole (each
// from two source dwords, coalesced over the workgroup), the partial dwords
// at its two ends as bytes (the neighbouring tiles write the rest of them).
constexpr int kTileCopyPer = 4;  // dwords per thread in flight

__global__ __launch_bounds__(256) void k_compact_tiles(const uint8_t* __restrict__ slots, uint32_t slot_bytes,
                                                       uint32_t tile_blocks, const uint64_t* __restrict__ boff,
                                                       uint64_t nblocks, uint8_t* __restrict__ dst) {
  const uint64_t t = blockIdx.x;
  const uint64_t b0 = t * tile_blocks, b1 = min(b0 + tile_blocks, nblocks);
  const uint64_t o = boff[b0], o1 = boff[b1];
  const uint8_t* s = slots + b0 * slot_bytes;
  // the tile's slots end at dword smax + 1; reads are clamped there, which
  // only an index built over rejected blocks (meta 0xFFFFFFFF, whose bytes
  // are meaningless anyway) can reach
  const uint64_t smax = (b1 - b0) * slot_bytes / 4 - 1;
  const uint32_t* s32 = reinterpret_cast<const uint32_t*>(s);
  uint32_t* d32 = reinterpret_cast<uint32_t*>(dst);
  // whole dwords [d0, d1) of dst inside [o, o1); dword d0 + k is source
  // bytes [in0 + 4k, in0 + 4k + 4)
  const uint64_t d0 = (o + 3) >> 2, d1 = o1 >> 2;
  const uint64_t ndw = d1 > d0 ? d1 - d0 : 0;
  const uint32_t in0 = (uint32_t)(4 * d0 - o);
  for (uint64_t k0 = threadIdx.x; k0 < ndw; k0 += 256 * kTileCopyPer) {
    uint32_t lo[kTileCopyPer], hi[kTileCopyPer];
#pragma unroll
    for (int u = 0; u < kTileCopyPer; ++u) {
      const uint64_t k = k0 + 256 * u;
      lo[u] = hi[u] = 0;
      if (k < ndw) {
        lo[u] = s32[min(k, smax)];
        hi[u] = s32[min(k + 1, smax)];
      }
    }
#pragma unroll
    for (int u = 0; u < kTileCopyPer; ++u) {
      const uint64_t k = k0 + 256 * u;
      if (k < ndw) d32[d0 + k] = __builtin_amdgcn_alignbyte(hi[u], lo[u], in0);
    }
  }
  // the bytes before d0 and from d1 on (at most 3 each; a tile holds at
  // least 4 bytes unless it is empty)
  const int x = threadIdx.x;
  if (x < 6 && o1 > o) {
    const uint64_t p = x < 3 ? o + x : 4 * d1 + (uint64_t)(x - 3);
    const uint64_t lim = x < 3 ? min(4 * d0, o1) : o1;
    if (p < lim && (x < 3 || p >= 4 * d0)) dst[p] = s[min(p - o, 4 * smax + 3)];
  }
}

__global__ void k_fill_lb(uint64_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Job table for streams cut into 128-value blocks: stream s covers values
// [val_off[s], val_off[s+1]) with flag flags[s]; its blocks are numbered from
// blk_off[s] (blk_off[nstreams] blocks in all, read on the device). Block
// k's stream is the last one starting at or before it (so empty streams are
// skipped). A wavefront takes runs of 1024 consecutive blocks, lane l the
// blocks base + 64q + l (stores coalesced over the wave): the first by
// binary search, each next one by a gallop up from the stream of the last.
constexpr uint64_t kJobRun = 64 * 16;

__global__ __launch_bounds__(256) void k_make_jobs(const uint64_t* __restrict__ val_off,
                                                   const uint64_t* __restrict__ blk_off,
                                                   const uint8_t* __restrict__ flags, uint64_t nstreams,
                                                   uint64_t* __restrict__ in_off, uint32_t* __restrict__ in_meta) {
  const uint64_t nblocks = blk_off[nstreams];
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = wave * kJobRun; base < nblocks; base += nwaves * kJobRun) {
    uint64_t k = base + (uint64_t)lane_id();
    if (k >= nblocks) continue;
    uint64_t lo = 0, hi = nstreams - 1;  // the last s with blk_off[s] <= k
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) >> 1;
      if (blk_off[mid] <= k)
        lo = mid;
      else
        hi = mid - 1;
    }
    uint64_t s = lo;
    for (int q = 0; q < (int)(kJobRun / 64) && k < nblocks; ++q, k += 64) {
      if (q > 0 && s + 1 < nstreams && blk_off[s + 1] <= k) {
        // gallop up to a bracket [s, t) with blk_off[s] <= k < blk_off[t]
        uint64_t t = s + 1, step = 1;
        while (t < nstreams && blk_off[t] <= k) {
          s = t;
          step <<= 1;
          t = s + step;
        }
        uint64_t h = t < nstreams ? t : nstreams;  // blk_off[h] > k or h == nstreams
        while (h - s > 1) {
          const uint64_t mid = (s + h) >> 1;
          if (blk_off[mid] <= k)
            s = mid;
          else
            h = mid;
        }
      }
      const uint64_t i = k - blk_off[s];
      const uint64_t v0 = val_off[s], nv = val_off[s + 1] - v0;
      const uint64_t rest = nv - 128 * i;
      in_off[k] = v0 + 128 * i;
      in_meta[k] = (uint32_t)(rest < 128 ? rest : 128) | ((uint32_t)flags[s] << 8);
    }
  }
}

}  // namespace utl

using namespace utl;

hipError_t utl_launch_encode_u32(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                 uint64_t nblocks, uint8_t* slots, uint32_t slot_bytes, uint32_t* out_meta,
                                 uint32_t* work_out, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = (uint64_t)kEncGrpWaves * kWaveBlocksEnc;
  const uint64_t grid = (nblocks + per_wg - 1) / per_wg;
  k_encode_u32<false><<<dim3((unsigned)grid), dim3(64 * kEncGrpWaves), 0, stream>>>(
      values, in_off, in_meta, nblocks, slots, slot_bytes, out_meta, work_out, nullptr, nullptr, 0);
  return hipGetLastError();
}

size_t utl_encode_direct_tmp_bytes(uint64_t nblocks) {
  const uint64_t ntiles = (nblocks + 31) / 32;
  return (size_t)((ntiles + 2) * sizeof(uint64_t));
}

hipError_t utl_launch_encode_u32_direct(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                        uint64_t nblocks, uint8_t* dst, uint64_t dst_cap, uint32_t* out_meta,
                                        uint64_t* block_off, uint32_t* work_out, uint64_t* lb, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = (uint64_t)kEncGrpWaves * kWaveBlocksEnc;
  const uint64_t ntiles = (nblocks + per_wg - 1) / per_wg;
  const uint64_t nlb = ntiles + 2;
  k_fill_lb<<<dim3((unsigned)((nlb + 255) / 256)), dim3(256), 0, stream>>>(lb, nlb);
  k_encode_u32<true><<<dim3((unsigned)ntiles), dim3(64 * kEncGrpWaves), 0, stream>>>(
      values, in_off, in_meta, nblocks, dst, 0, out_meta, work_out, lb, block_off, dst_cap);
  return hipGetLastError();
}

hipError_t utl_launch_encode_u32_size(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                      uint64_t nblocks, uint32_t* out_meta, uint64_t* plan, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = 4 * (uint64_t)kWaveBlocksEnc;
  k_encode_u32_size<<<dim3((unsigned)((nblocks + per_wg - 1) / per_wg)), dim3(256), 0, stream>>>(
      values, in_off, in_meta, nblocks, out_meta, reinterpret_cast<uint2*>(plan));
  return hipGetLastError();
}

hipError_t utl_launch_encode_u32_write(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                       uint64_t nblocks, const uint64_t* plan, const uint64_t* block_off, uint8_t* dst,
                                       uint64_t dst_cap, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = (uint64_t)kEncGrpWaves * kWaveBlocksEnc;
  k_encode_u32_write<<<dim3((unsigned)((nblocks + per_wg - 1) / per_wg)), dim3(64 * kEncGrpWaves), 0, stream>>>(
      values, in_off, in_meta, nblocks, reinterpret_cast<const uint2*>(plan), block_off, dst, dst_cap);
  return hipGetLastError();
}

hipError_t utl_launch_encode_u64(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                 uint64_t nblocks, uint8_t* slots, uint32_t slot_bytes, uint32_t* out_meta,
                                 hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = (uint64_t)kEncGrpWaves * kWaveBlocksEnc;
  const uint64_t grid = (nblocks + per_wg - 1) / per_wg;
  k_encode_u64<kEncSlots><<<dim3((unsigned)grid), dim3(64 * kEncGrpWaves), 0, stream>>>(
      values, in_off, in_meta, nblocks, slots, slot_bytes, nullptr, 0, out_meta);
  return hipGetLastError();
}

hipError_t utl_launch_encode_u64_size(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                      uint64_t nblocks, uint32_t* out_meta, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = (uint64_t)kEncGrpWaves * kWaveBlocksEnc;
  k_encode_u64<kEncSize><<<dim3((unsigned)((nblocks + per_wg - 1) / per_wg)), dim3(64 * kEncGrpWaves), 0, stream>>>(
      values, in_off, in_meta, nblocks, nullptr, 0, nullptr, 0, out_meta);
  return hipGetLastError();
}

hipError_t utl_launch_encode_u64_write(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                       uint64_t nblocks, const uint64_t* block_off, uint8_t* dst, uint64_t dst_cap,
                                       hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t per_wg = (uint64_t)kEncGrpWaves * kWaveBlocksEnc;
  k_encode_u64<kEncWrite><<<dim3((unsigned)((nblocks + per_wg - 1) / per_wg)), dim3(64 * kEncGrpWaves), 0, stream>>>(
      values, in_off, in_meta, nblocks, dst, 0, block_off, dst_cap, nullptr);
  return hipGetLastError();
}

hipError_t utl_launch_compact(const uint8_t* slots, uint32_t slot_bytes, bool tiles, const uint64_t* boff,
                              uint64_t nblocks, uint8_t* dst, hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  if (tiles) {
    const uint32_t per_tile = kEncGrpWaves * kWaveBlocksEnc;
    const uint64_t ntiles = (nblocks + per_tile - 1) / per_tile;
    k_compact_tiles<<<dim3((unsigned)ntiles), dim3(256), 0, stream>>>(slots, slot_bytes, per_tile, boff, nblocks,
                                                                      dst);
  } else {
    k_compact<<<dim3((unsigned)((nblocks + 31) / 32)), dim3(256), 0, stream>>>(slots, slot_bytes, boff, nblocks, dst);
  }
  return hipGetLastError();
}

hipError_t utl_launch_make_jobs(const uint64_t* val_off, const uint64_t* blk_off, const uint8_t* flags,
                                uint64_t nstreams, uint64_t* in_off, uint32_t* in_meta, hipStream_t stream) {
  if (nstreams == 0) return hipSuccess;
  k_make_jobs<<<dim3(2048), dim3(256), 0, stream>>>(val_off, blk_off, flags, nstreams, in_off, in_meta);
  return hipGetLastError();
}
