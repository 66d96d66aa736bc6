// utl_decode.hip -- gfx950 decode kernels for UTL-PFOR blocks.
//
// Eight threads decode one block: thread j (0..7) of a block owns the UTL
// lane pair (2j, 2j+1), i.e. the 16 values at positions 16r + 2j + e. A
// wavefront decodes 8 consecutive blocks and a 256-thread workgroup a tile
// of 32. Tiles are assigned XCD-major (workgroup g runs on XCD g % 8), so
// each XCD's L2 sees one contiguous range of the stream and cache lines
// shared by neighbouring blocks are fetched from HBM once.
//
// A wavefront first stages the contiguous bytes of its blocks into an LDS
// arena with 16-byte loads of the aligned chunks that hold them (one round
// trip to HBM; groups too large for the arena take further rounds), then
// decodes every block from LDS following unpackUint32Scalar
// (scalar.go:260-309) and its validation order (unpack_block.go:127-237):
//   validate -> unpack (one specialised path per bit width) -> exceptions
//   (the StreamVByte values of all eight blocks decoded by the whole wave
//   from one flat table of control bytes; bitmap patch by popcount rank,
//   position-list patch by LDS OR) -> per-lane delta in registers (zigzag,
//   or overflow checked, first wrap in v-major order) -> + FOR base ->
//   8-byte stores of the lane pairs, or, for rounds with blocks at odd
//   offsets or shorter than 128 values, 16-byte stores of the round's whole
//   output range staged in LDS.
#include "utl_device.h"
#include "utl_format.h"
#include "utl_kernels.h"

namespace utl {

constexpr int kTpb = 8;             // threads per block
constexpr int kWaveBlocks = 8;      // blocks per wavefront
// Bytes of a sub-block the decoder can read: header, svbLen, FOR64 base,
// block2Len (16) + payload (512) + bitmap (16) + control (32) + data (512).
constexpr int kSubExtent = 1088;
// Staged bytes per wavefront (the arena) and wavefronts per workgroup are
// launch variants (template parameters), chosen per workload by the
// autotuner. u32 arenas: 2560 B (24 waves per CU by LDS; a wave whose eight
// blocks span more takes further staging rounds, each a round trip to HBM),
// 3328 B (20 or 22 waves per CU) and 4352 B (16 or 18), which holds eight
// blocks of bit width 32 (8 x 528 + 19 bytes) in one round. u64 blocks use
// 8320 B, which holds all eight blocks of a wave in one round even when they
// are two-block blocks of bw 32 (7 * 1034 + 1034 + 19 bytes), at three
// workgroups per CU (two staging round trips per wave cost more: 8Mi uniform
// two-block 30.5 -> 24.7 us).
constexpr int kArena32Small = 2560;
constexpr int kArena32Mid = 3328;
constexpr int kArena32Large = 4352;
constexpr int kArena64 = 8320;
// Exceptions of the wave's blocks are numbered flat: block b's control bytes
// are control bytes [cbase_b, cbase_b + ctrl_b) of the wave, with cbase_b the
// control bytes of the blocks before it, and its exception i is slot
// 4*cbase_b + i of aux. The table entry of control byte c waits in slot
// kTabBase + c (consecutive entries: no LDS bank conflicts). Round k of the
// decode loop reads entries [64k, 64k + 64) and then writes the values of
// those control bytes to slots [256k, 256k + 256); those reach kTabBase + c
// only for entries this round or an earlier one has already read, so the
// table can share the slots. A block in position-list mode later reuses
// aux[128*b ..) to scatter.
constexpr int kAux = kWaveBlocks * kBlock + 4;  // flat exception slots (+ room for reads past the last)
constexpr int kCtrlMax = kWaveBlocks * 32;       // control bytes of a wave's blocks
constexpr int kTabBase = kWaveBlocks * kBlock - kCtrlMax;  // 768
static_assert(4 * 64 * (kCtrlMax / 64 - 1) + 4 * 64 <= kTabBase + kCtrlMax, "table rounds");

// A table entry holds a data offset into the arena (14 bits: every arena is
// below 16 KiB), the block's bit width at bit 14 and the control byte at 24.
constexpr uint32_t kTabOff = 0x3FFF;
constexpr int kTabBwShift = 14;

template <int ARENA>
struct alignas(16) WaveLdsT {
  static_assert(ARENA + 64 <= (int)kTabOff + 1, "table offsets cover the arena");
  static constexpr int kArenaBytes = ARENA;
  uint32_t arena[ARENA / 4];
  uint32_t aux[kAux];  // exception high bits by flat slot (and the table entries before that)
};
// LDS per wavefront: arena + 4112 B of exception slots. u32 with the small
// arena: 6672 B, six 4-wave workgroups per CU; u64: 12432 B, three.
static_assert(sizeof(WaveLdsT<kArena32Small>) == 6672, "LDS per wavefront");
static_assert(6 * 4 * sizeof(WaveLdsT<kArena32Small>) + 8 <= 160 * 1024, "six u32 workgroups per CU");
static_assert(5 * 4 * sizeof(WaveLdsT<kArena32Mid>) + 8 <= 160 * 1024, "five u32 workgroups per CU");
static_assert(9 * 2 * sizeof(WaveLdsT<kArena32Large>) + 8 <= 160 * 1024, "nine 2-wave u32 workgroups per CU");
static_assert(3 * 4 * sizeof(WaveLdsT<kArena64>) + 8 <= 160 * 1024, "three u64 workgroups per CU");

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP row_shr:n within 16-lane rows, masked to 8-lane groups.
template <int N>
__device__ __forceinline__ uint32_t grp_shr(uint32_t x, int j) {
  const uint32_t y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + N, 0xF, 0xF, true);
  return j >= N ? y : 0u;
}

__device__ __forceinline__ uint32_t grp_incl_scan(uint32_t x, int j) {
  x += grp_shr<1>(x, j);
  x += grp_shr<2>(x, j);
  x += grp_shr<4>(x, j);
  return x;
}

__device__ __forceinline__ uint32_t grp_last(uint32_t x) { return (uint32_t)__shfl((int)x, lane_id() | 7, 64); }

__device__ __forceinline__ uint32_t grp_min(uint32_t x) {
  x = min(x, (uint32_t)__shfl_xor((int)x, 1, 64));
  x = min(x, (uint32_t)__shfl_xor((int)x, 2, 64));
  x = min(x, (uint32_t)__shfl_xor((int)x, 4, 64));
  return x;
}

// n (1..4) little-endian bytes at LDS byte offset off.
__device__ __forceinline__ uint32_t lds_bytes(const uint32_t* a, int off, int n) {
  const int q = off >> 2, sh = off & 3;
  const uint32_t v = __builtin_amdgcn_alignbyte(a[q + 1], a[q], sh);
  return n >= 4 ? v : (v & ((1u << (8 * n)) - 1u));
}
__device__ __forceinline__ uint32_t lds_u32(const uint32_t* a, int off) {
  const int q = off >> 2, sh = off & 3;
  return __builtin_amdgcn_alignbyte(a[q + 1], a[q], sh);
}

// UTL unpack of the lane pair (2j, 2j+1) for a fixed bit width
// (scalar.go:65-116): lane L's word w is at payload byte 64w + 4L; value r
// occupies bits [r*BW, r*BW+BW) of the lane's bit stream.
template <int BW>
__device__ __forceinline__ void unpack_bw(const uint32_t* a, int pbyte, int j, uint32_t* v0, uint32_t* v1) {
  if constexpr (BW == 0) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      v0[r] = 0;  // (separate stores: a chained assignment here sends both arrays to scratch)
      v1[r] = 0;
    }
  } else {
    constexpr int NW = BW / 4;
    uint32_t w0[NW + 1], w1[NW + 1];
    const int base = pbyte + 8 * j;
    const int q = base >> 2, sh = base & 3;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const uint32_t d0 = a[q + 16 * k], d1 = a[q + 16 * k + 1], d2 = a[q + 16 * k + 2];
      w0[k] = __builtin_amdgcn_alignbyte(d1, d0, sh);
      w1[k] = __builtin_amdgcn_alignbyte(d2, d1, sh);
    }
    w0[NW] = w1[NW] = 0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int bit = r * BW, wi = bit >> 5, o = bit & 31;
      if constexpr (BW == 32) {
        v0[r] = w0[wi];
        v1[r] = w1[wi];
      } else {
        constexpr uint32_t m = (1u << BW) - 1u;
        if (o + BW <= 32) {
          v0[r] = (w0[wi] >> o) & m;
          v1[r] = (w1[wi] >> o) & m;
        } else {
          v0[r] = __builtin_amdgcn_alignbit(w0[wi + 1], w0[wi], o) & m;
          v1[r] = __builtin_amdgcn_alignbit(w1[wi + 1], w1[wi], o) & m;
        }
      }
    }
  }
}

__device__ __forceinline__ void unpack_any(int bw, const uint32_t* a, int pbyte, int j, uint32_t* v0, uint32_t* v1) {
  switch (bw) {
    case 0: unpack_bw<0>(a, pbyte, j, v0, v1); break;
    case 4: unpack_bw<4>(a, pbyte, j, v0, v1); break;
    case 8: unpack_bw<8>(a, pbyte, j, v0, v1); break;
    case 12: unpack_bw<12>(a, pbyte, j, v0, v1); break;
    case 16: unpack_bw<16>(a, pbyte, j, v0, v1); break;
    case 20: unpack_bw<20>(a, pbyte, j, v0, v1); break;
    case 24: unpack_bw<24>(a, pbyte, j, v0, v1); break;
    case 28: unpack_bw<28>(a, pbyte, j, v0, v1); break;
    default: unpack_bw<32>(a, pbyte, j, v0, v1); break;
  }
}

struct SubResult {
  int count;
  int consumed;
};

// Layout of one sub-block after the header checks of unpack_block.go:62-172.
struct SubPlan {
  Hdr h;
  bool u64s;       // FOR64 single block (its base is applied by the caller)
  int poff;        // payload offset from the sub-block's first byte
  int svb_len;
  uint32_t for_base;
};

// Header checks up to the payload (unpack_block.go:62-172), without branches:
// every field is read (all reads stay within the staged bytes of the block
// and the arena) and the status is chosen by the reference's precedence:
// length < 4, integer type, count 0 (a valid empty block of 4 bytes), count
// > 128, bw > 32, short for payload + index, short for svbLen. Returns the
// status; go is set when the sub-block has values to decode.
template <bool ForU64>
__device__ __forceinline__ int32_t sub_prelude(const uint32_t* a, int at, int64_t len, SubPlan& p, bool& go,
                                               int& consumed) {
  p.h = decode_header(lds_u32(a, at));
  const Hdr& h = p.h;
  const bool combine = ForU64 && h.combine;  // the u32 API ignores it (unpack_block.go:81)
  const bool type_ok = ForU64 ? h.int_type == 3 : (h.int_type == 2 || h.int_type == 1);
  const bool has_exc = h.exc > 0;
  const bool has_for = h.for_width > 0;
  p.u64s = ForU64 && is_for64_single(h.int_type, h.for_width, combine);
  const int for_off = kHdr + (has_exc ? kSvbLen : 0);
  const int fbb = for_base_bytes(h.for_width);
  const int poff = for_off + (has_for ? (p.u64s ? kFor64 : fbb) : 0) + (combine ? kB2Len : 0);
  p.poff = poff;
  const int64_t end = (int64_t)poff + h.bw * 16 + (has_exc ? exc_index_size(h.exc) : 0);
  const uint32_t svb_raw = lds_bytes(a, at + kHdr, 2);
  const uint32_t for_raw = lds_u32(a, at + for_off);
  p.svb_len = has_exc ? (int)svb_raw : 0;
  p.for_base = (has_for && !p.u64s) ? (fbb >= 4 ? for_raw : (for_raw & ((1u << (8 * fbb)) - 1u))) : 0u;
  int32_t st = (has_exc && end + p.svb_len > len) ? UTL_ERR_INVALID_BUFFER : UTL_OK;
  st = len < end ? UTL_ERR_INVALID_BUFFER : st;
  st = (unsigned)h.bw > 32u ? UTL_ERR_INVALID_BUFFER : st;
  st = h.count > kBlock ? UTL_ERR_INVALID_BLOCK_LENGTH : st;
  st = h.count == 0 ? UTL_OK : st;
  st = type_ok ? st : UTL_ERR_UNSUPPORTED_TYPE;
  st = len < kHdr ? UTL_ERR_INVALID_BUFFER : st;
  go = st == UTL_OK && h.count > 0;
  consumed = st == UTL_OK && h.count == 0 ? kHdr : 0;
  return st;
}

// Bitmap patch of one value pair at bits b, b+1 of bitmap word w, whose
// lower bits hold `rank` set bits: x[rank] is the high part of the first set
// one (exceptions.go:246-284).
__device__ __forceinline__ void patch_pair(const uint32_t* x, uint32_t w, int b, uint32_t rank, uint32_t& va,
                                           uint32_t& vb) {
  const uint32_t lo = x[rank], hi = x[rank + 1];
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_sbfe((int)w, b, 1);      // 0 or ~0
  const uint32_t s1 = (uint32_t)__builtin_amdgcn_sbfe((int)w, b + 1, 1);
  va |= lo & s0;
  vb |= ((hi & s0) | (lo & ~s0)) & s1;
}

// One StreamVByte value (svb.go:9-24, exceptions.go:32-88) of code `code`
// (code + 1 bytes) at arena byte off, shifted into the high bits by bw
// (exceptions.go:83-86; a shift by 32 gives 0).
__device__ __forceinline__ uint32_t svb_at(const uint32_t* a, int off, uint32_t code, int bw) {
  const int q = off >> 2;
  uint32_t v = __builtin_amdgcn_alignbyte(a[q + 1], a[q], off);  // uses off & 3
  const uint32_t drop = (code ^ 3) << 3;                          // bytes past the value's length
  v = (v << drop) >> drop;
  return (uint32_t)((uint64_t)v << bw);
}

// Decodes one staged sub-block per 8-lane group whose first byte is at arena
// byte `at` and whose true length is len; groups with act == false only take
// part in the wave-wide steps. blk is the group's index in the wave, j the
// thread's index in the group. Returns the status word; count and consumed
// are zero unless the status is UTL_OK. Every lane of the wave must call it.
template <bool ForU64, class WL>
__device__ __forceinline__ int32_t decode_sub(WL* W, bool act, int blk, int at, int64_t len, int j,
                                              uint32_t* v0, uint32_t* v1, SubResult& res) {
  const uint32_t* a = W->arena;
  // Every lane parses a header (lanes not in the round read the arena's
  // first bytes: callers pass at = 0), so there is no branch and no default
  // plan to set up; those lanes then take part in nothing.
  SubPlan p;
  bool go;
  int consumed;
  int32_t st = sub_prelude<ForU64>(a, at, len, p, go, consumed);
  if (!act) {
    st = UTL_OK;
    go = false;
    consumed = 0;
  }
  res.count = 0;
  res.consumed = 0;
  if (go) unpack_any(p.h.bw, a, at + p.poff, j, v0, v1);

  // ---- exception checks and this block's control-byte table (exceptions.go:32-88) ----
  // Computed without branches by every thread; a group contributes control
  // bytes to the wave table (ctrl_n) only when all its checks pass. The
  // three checks (excCount > count, control bytes > svbLen, control sums >
  // svbLen) all give ErrInvalidBuffer, so their order does not matter.
  const int exc = go ? p.h.exc : 0;
  const int exc_start = at + p.poff + p.h.bw * 16;
  if (go) consumed = p.poff + p.h.bw * 16;
  const int ctrl = svb_ctrl_count(exc);
  const int svb_start = exc_start + exc_index_size(exc);
  // Thread j owns control bytes 4j .. 4j+3 (exceptions 16j .. 16j+15). Only
  // the slots of exceptions < excCount count (exceptions.go:63-80): data
  // bytes = valid slots + sum of their 2-bit codes, summed per byte with
  // SWAR arithmetic.
  const int nv = min(max(exc - 16 * j, 0), 16);  // valid slots in this dword
  const uint32_t cw = lds_u32(a, nv > 0 ? svb_start + 4 * j : 0);
  const uint32_t cm = cw & (nv >= 16 ? 0xFFFFFFFFu : ((1u << (2 * nv)) - 1u));
  const uint32_t t2 = (cm & 0x33333333u) + ((cm >> 2) & 0x33333333u);
  const uint32_t t4 = (t2 & 0x0F0F0F0Fu) + ((t2 >> 4) & 0x0F0F0F0Fu);  // code sum per byte
  const uint32_t pre = t4 * 0x01010100u;                                // exclusive byte prefix
  const int part = nv + (int)((t4 * 0x01010101u) >> 24);
  const uint32_t incl = grp_incl_scan((uint32_t)part, j);
  const int total = (int)grp_last(incl);
  const bool svb_ok = exc <= p.h.count && ctrl + total <= p.svb_len;
  if (exc > 0 && !svb_ok) st = UTL_ERR_INVALID_BUFFER;
  const bool has_svb = exc > 0 && svb_ok;
  const int ctrl_n = has_svb ? ctrl : 0;  // control bytes this group contributes to the wave table
  if (has_svb) consumed = svb_start + p.svb_len - at;
  // Data offset of each owned control byte, with its bw and the byte.
  const uint32_t off = (uint32_t)(svb_start + ctrl + (int)incl - part);
  const uint32_t bwf = (uint32_t)p.h.bw << kTabBwShift;
  uint4 te;
  te.x = (off + ((pre >> 0) & 0xFF)) | bwf | ((cw & 0xFFu) << 24);
  te.y = (off + min(4, nv) + ((pre >> 8) & 0xFF)) | bwf | (((cw >> 8) & 0xFFu) << 24);
  te.z = (off + min(8, nv) + ((pre >> 16) & 0xFF)) | bwf | (((cw >> 16) & 0xFFu) << 24);
  te.w = (off + min(12, nv) + (pre >> 24)) | bwf | ((cw >> 24) << 24);

  // ---- wave-wide StreamVByte decode of every group's exceptions ----
  const uint32_t cincl = wave_group_scan((uint32_t)ctrl_n);
  const int cbase = (int)cincl - ctrl_n;
  const int ctot = __builtin_amdgcn_readlane((int)cincl, 63);
  if (ctot > 0) {
    if (ctrl_n > 0) {
      const int c0 = 4 * j;
      uint32_t* t = &W->aux[kTabBase + cbase + c0];  // entry of control byte c in slot kTabBase + c
      if (c0 < ctrl_n) t[0] = te.x;
      if (c0 + 1 < ctrl_n) t[1] = te.y;
      if (c0 + 2 < ctrl_n) t[2] = te.z;
      if (c0 + 3 < ctrl_n) t[3] = te.w;
    }
    wave_fence();
    // Lane c decodes the four values of control byte c (its table entry in
    // slot kTabBase + c) into slots 4c .. 4c+3. Slots past a block's
    // excCount decode bytes nobody reads.
#pragma unroll 1
    for (int c = lane_id(); c < ctot; c += 64) {
      const uint32_t e = W->aux[kTabBase + c];
      const uint32_t codes = e >> 24;
      const int bw = (int)((e >> kTabBwShift) & 63);
      const int o0 = (int)(e & kTabOff);
      const int o1 = o0 + (int)(codes & 3u) + 1;
      const int o2 = o1 + (int)((codes >> 2) & 3u) + 1;
      const int o3 = o2 + (int)((codes >> 4) & 3u) + 1;
      uint4 x;
      x.x = svb_at(a, o0, codes & 3u, bw);
      x.y = svb_at(a, o1, (codes >> 2) & 3u, bw);
      x.z = svb_at(a, o2, (codes >> 4) & 3u, bw);
      x.w = svb_at(a, o3, codes >> 6, bw);
      *reinterpret_cast<uint4*>(&W->aux[4 * c]) = x;
    }
    wave_fence();
  }

  // ---- patch (exceptions.go:194-284) ----
  const bool patch = ctrl_n > 0;
  const bool list = patch && p.h.exc <= kExcBitmap;
  uint32_t hv0 = 0, hv1 = 0;
  if (patch) {
    const uint32_t* x = &W->aux[4 * cbase];
    if (list) {
      hv0 = j < p.h.exc ? x[j] : 0u;
      hv1 = j + 8 < p.h.exc ? x[j + 8] : 0u;
    } else {
      // 128-bit bitmap masked to count; the rank of a set bit is the index
      // of its high bits. popcount(bitmap) > excCount applies nothing.
      // the 16 bitmap bytes: the five dwords that hold them, read from one
      // address, aligned in registers
      uint32_t bm[4];
      {
        const uint32_t* q = a + (exc_start >> 2);
        const int sh = exc_start & 3;
        uint32_t w[5];
#pragma unroll
        for (int d = 0; d < 5; ++d) w[d] = q[d];
#pragma unroll
        for (int d = 0; d < 4; ++d) bm[d] = __builtin_amdgcn_alignbyte(w[d + 1], w[d], sh);
      }
      if (p.h.count < kBlock) {
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int keep = p.h.count - 32 * d;
          bm[d] = keep >= 32 ? bm[d] : (keep <= 0 ? 0u : (bm[d] & ((1u << keep) - 1u)));
        }
      }
      const uint32_t pre1 = __popc(bm[0]), pre2 = pre1 + __popc(bm[1]), pre3 = pre2 + __popc(bm[2]);
      if ((int)(pre3 + __popc(bm[3])) <= p.h.exc) {
        // word d holds rows 2d (bits 2j, 2j+1) and 2d+1 (bits 16+2j, 17+2j)
        const int ba = 2 * j, bb = 16 + 2 * j;
        const uint32_t ma = (1u << ba) - 1u, mb = (1u << bb) - 1u;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const uint32_t w = bm[d];
          const uint32_t pd = d == 0 ? 0u : d == 1 ? pre1 : d == 2 ? pre2 : pre3;
          patch_pair(x, w, ba, pd + __popc(w & ma), v0[2 * d], v1[2 * d]);
          patch_pair(x, w, bb, pd + __popc(w & mb), v0[2 * d + 1], v1[2 * d + 1]);
        }
      }
    }
  }
  // Sorted position list: OR each exception into its position in LDS. Full
  // blocks mask positions with 0x7F, partial blocks skip positions >= count
  // (exceptions.go:211-230). Every bitmap read of aux is done before the
  // scatter reuses aux[128*blk ..).
  if (__ballot(list)) {
    wave_fence();
    uint32_t* y = &W->aux[kBlock * blk];
    if (list) {
#pragma unroll
      for (int r = 0; r < kRows; ++r) *reinterpret_cast<uint2*>(&y[16 * r + 2 * j]) = make_uint2(v0[r], v1[r]);
    }
    wave_fence();
    if (list) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int i = j + 8 * e;
        if (i < p.h.exc) {
          const uint32_t q = lds_bytes(a, exc_start + i, 1);
          const uint32_t hv = e ? hv1 : hv0;
          if (p.h.count == kBlock) atomicOr(&y[q & 0x7F], hv);
          else if ((int)q < p.h.count) atomicOr(&y[q], hv);
        }
      }
    }
    wave_fence();
    if (list) {
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const uint2 t = *reinterpret_cast<const uint2*>(&y[16 * r + 2 * j]);
        v0[r] = t.x;
        v1[r] = t.y;
      }
    }
  }

  if (!go) res.consumed = st == UTL_OK ? consumed : 0;  // count 0: a valid 4-byte block

  // ---- per-lane delta decode (delta.go:50-115) and FOR (for.go:136-140) ----
  // (One exit for every path: early returns here cost register copies on
  // the common path.)
  if (go && st == UTL_OK) {
    const uint32_t fb = (p.h.for_width > 0 && !p.u64s) ? p.for_base : 0u;
    if (p.h.delta) {
      // Zigzag addends are decoded in place; their running sums wrap freely
      // (delta.go:50-72). The prefix runs in place for every block, with the
      // base folded into row 0. Other blocks report the first wrap of the
      // running sum (delta.go:83-115), which needs an addend >= 2^29 in the
      // lane: only waves holding one look for it, on the sums less the base.
      // (No path rewrites the values: branches that did cost register
      // copies on the common path.)
      bool check = false;
      if (p.h.zigzag) {
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          v0[r] = zz_dec(v0[r]);
          v1[r] = zz_dec(v1[r]);
        }
      } else {
        uint32_t big = 0;
#pragma unroll
        for (int r = 0; r < kRows; ++r) big |= v0[r] | v1[r];
        check = big >= (1u << 29);
      }
      v0[0] += fb;
      v1[0] += fb;
#pragma unroll
      for (int r = 1; r < kRows; ++r) {
        v0[r] += v0[r - 1];
        v1[r] += v1[r - 1];
      }
      if (__ballot(check) != 0) {
        // A wrap at row r leaves sum[r] < sum[r-1] (an addend is < 2^32).
        // The first wrap at a position < count in v-major order wins.
        uint32_t first = 0xFFFFFFFFu;
        if (!p.h.zigzag) {
#pragma unroll
          for (int r = kRows - 1; r >= 1; --r) {
            const uint32_t p0 = 16 * r + 2 * j, p1 = p0 + 1;
            if (v1[r] - fb < v1[r - 1] - fb && (int)p1 < p.h.count) first = p1;
            if (v0[r] - fb < v0[r - 1] - fb && (int)p0 < p.h.count) first = p0;
          }
        }
        const uint32_t m = grp_min(first);
        if (m != 0xFFFFFFFFu) st = UTL_ERR_OVERFLOW | (int32_t)(m << 8);
      }
    } else if (fb != 0) {
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        v0[r] += fb;
        v1[r] += fb;
      }
    }
    if (st == UTL_OK) {
      res.count = p.h.count;
      res.consumed = consumed;
    }
  }
  return st;
}

// Staged extent of a block: u32 blocks read at most kSubExtent bytes; a
// uint64 block's two sub-blocks together at most 2 * 1018 + 2 when valid.
constexpr int kExtent64 = 2176;

template <bool U64>
__device__ __forceinline__ int64_t clamp_extent(int64_t len) {
  constexpr int64_t kExt = U64 ? kExtent64 : kSubExtent;
  return len < 0 ? 0 : (len < kExt ? len : kExt);
}

// Copies bytes [g, g+n) into the arena at byte (g & 15): every 16-byte
// aligned chunk that holds one of the bytes, spread over the wave.
// All of a thread's chunk loads are issued before the first LDS write, so a
// staged range costs one round trip to HBM.
// The chunks of one staged range, loaded into registers (scalar members: a
// uint4 array here is placed in scratch).
template <int N>
struct Stage {
  uint4 c[N];
  int nchunks;
};

template <class WL>
__device__ __forceinline__ void stage_range(WL* W, const uint8_t* __restrict__ g, int64_t n) {
  constexpr int N = (WL::kArenaBytes / 16 + 63) / 64;  // chunks per thread
  const int s = (int)(reinterpret_cast<uintptr_t>(g) & 15);
  const uint4* base = reinterpret_cast<const uint4*>(g - s);
  Stage<N> st;
  st.nchunks = n > 0 ? (int)((s + n + 15) >> 4) : 0;
  const int l = lane_id();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    st.c[k] = make_uint4(0, 0, 0, 0);
    if (l + 64 * k < st.nchunks) st.c[k] = base[l + 64 * k];
  }
  uint4* ar = reinterpret_cast<uint4*>(W->arena);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (l + 64 * k < st.nchunks) ar[l + 64 * k] = st.c[k];
  wave_fence();
}

// Decoded values are written once and not read again by this kernel. A
// full block at a value offset that is a multiple of 16 writes whole 64-byte
// segments, one per row, which go out with non-temporal stores: they stream
// to HBM instead of allocating L2 lines (on the 64 Mi delta stream this took
// the kernel from 98 to 72 us). Any other block (partial, or at an offset
// that splits its rows across segments, as in streams of arbitrary lengths)
// goes through the wave's staged store below when it can; here it writes
// 4-byte pieces, which must be merged into lines in L2 first: non-temporal
// ones would reach HBM as partial lines.
__device__ __forceinline__ bool direct_store(int count, uint64_t out0) { return count == kBlock && (out0 & 15) == 0; }

__device__ __forceinline__ void store_u32(uint32_t* __restrict__ out, uint64_t out0, int count, int j,
                                          const uint32_t* v0, const uint32_t* v1) {
  uint32_t* o = out + out0 + 2 * j;
  if (direct_store(count, out0)) {
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      __builtin_nontemporal_store(((unsigned long long)v1[r] << 32) | v0[r],
                                  reinterpret_cast<unsigned long long*>(o + 16 * r));
  } else {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int p = 16 * r + 2 * j;
      if (p < count) o[16 * r] = v0[r];
      if (p + 1 < count) o[16 * r + 1] = v1[r];
    }
  }
}

// The values of one round whose blocks fill the output range [g0, g1)
// without holes (every block decoded, each filling its whole slot), staged
// in aux in output order and written out by the wave as 16-byte pieces of
// the range: every piece inside the range non-temporal, the values of the
// partial pieces at its two ends as cached 4-byte stores. Blocks at odd
// offsets or shorter than 128 values (streams of arbitrary lengths) then
// still write whole 16-byte pieces.
// aux is free: every exception read of the round is done. act: the lane's
// block is in the round and has count values at out0.
template <class WL>
__device__ __forceinline__ void store_u32_staged(WL* W, uint32_t* __restrict__ out, uint64_t g0, uint64_t g1,
                                                 bool act, uint64_t out0, int count, int j, const uint32_t* v0,
                                                 const uint32_t* v1) {
  uint32_t* y = W->aux;
  const int sh = (int)((reinterpret_cast<uintptr_t>(out + g0) >> 2) & 3);  // values before g0 in its 16 bytes
  const int n = (int)(g1 - g0);                                            // <= 8 * 128
  wave_fence();
  if (act) {
    uint32_t* z = y + sh + (int)(out0 - g0) + 2 * j;
    if (count == kBlock) {
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        z[16 * r] = v0[r];
        z[16 * r + 1] = v1[r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int p = 16 * r + 2 * j;
        if (p < count) z[16 * r] = v0[r];
        if (p + 1 < count) z[16 * r + 1] = v1[r];
      }
    }
  }
  wave_fence();
  // piece k: values [4k - sh, 4k - sh + 4) of the range, staged at y[4k]
  uint32_t* const ob = out + g0 - sh;  // 16-byte aligned
  const int npieces = (sh + n + 3) >> 2;  // <= 257: y[4k + 3] < kAux
#pragma unroll 1
  for (int k = lane_id(); k < npieces; k += 64) {
    const uint4 q = *reinterpret_cast<const uint4*>(&y[4 * k]);
    const int v = 4 * k - sh;  // first value of the piece, relative to g0
    uint4* o = reinterpret_cast<uint4*>(ob + 4 * k);
    // every whole piece non-temporal (measured on 200,000 uneven streams:
    // 4.03 ms, against 4.06 when only the pieces of segments wholly inside
    // the range and in one store instruction were, and 4.46 with none)
    if (v >= 0 && v + 4 <= n) {
      typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
      u32x4 t;
      t.x = q.x;
      t.y = q.y;
      t.z = q.z;
      t.w = q.w;
      __builtin_nontemporal_store(t, reinterpret_cast<u32x4*>(o));
    } else {
      uint32_t* e = ob + 4 * k;
      if (v >= 0 && v < n) e[0] = q.x;
      if (v + 1 >= 0 && v + 1 < n) e[1] = q.y;
      if (v + 2 >= 0 && v + 2 < n) e[2] = q.z;
      if (v + 3 < n) e[3] = q.w;  // v + 3 >= 0 always: v >= -3
    }
  }
  wave_fence();
}

// two: values are (hi << 32) | lo; else lo + base.
__device__ __forceinline__ void store_u64(uint64_t* __restrict__ out, uint64_t out0, int count, int j,
                                          const uint32_t* l0, const uint32_t* l1, const uint32_t* h0,
                                          const uint32_t* h1, bool two, uint64_t base) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  unsigned long long* o = reinterpret_cast<unsigned long long*>(out + out0 + 2 * j);
  const bool full = count == kBlock && (out0 & 1) == 0;  // 16-byte aligned pairs
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const unsigned long long y0 = two ? (((uint64_t)h0[r] << 32) | l0[r]) : (uint64_t)l0[r] + base;
    const unsigned long long y1 = two ? (((uint64_t)h1[r] << 32) | l1[r]) : (uint64_t)l1[r] + base;
    const int p = 16 * r + 2 * j;
    if (full) {
      u64x2 q;
      q.x = y0;
      q.y = y1;
      __builtin_nontemporal_store(q, reinterpret_cast<u64x2*>(o + 16 * r));
    } else {
      if (p < count) o[16 * r] = y0;
      if (p + 1 < count) o[16 * r + 1] = y1;
    }
  }
}

// uint64.go:191-224 unpackUint64Block up to its first sub-block: length and
// integer type, and the FOR64 base of a single block. Returns the status;
// go is set when the first sub-block is to be decoded.
__device__ __forceinline__ int32_t u64_prelude(const uint32_t* a, bool act, int at, int64_t len, Hdr& h,
                                               uint64_t& base, bool& go) {
  h = decode_header(0);
  base = 0;
  go = false;
  if (!act) return UTL_OK;
  if (len < kHdr) return UTL_ERR_INVALID_BUFFER;
  h = decode_header(lds_u32(a, at));
  if (h.int_type != 3) return UTL_ERR_UNSUPPORTED_TYPE;
  if (!h.combine && is_for64_single(h.int_type, h.for_width, h.combine)) {
    const int bo = kHdr + (h.exc > 0 ? kSvbLen : 0);
    if (len < bo + kFor64) return UTL_ERR_INVALID_BUFFER;
    base = (uint64_t)lds_u32(a, at + bo) | ((uint64_t)lds_u32(a, at + bo + 4) << 32);
  }
  go = true;
  return UTL_OK;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// Where a launch finds its blocks. Block form: block k is src[boff[k] :
// boff[k+1]] and decodes to out[ooff[k] ..] (the index scan's output, or a
// gathered index). Meta form (META): blocks are concatenated in order, block
// k's meta word (byte length | count << 24) is meta[k], and wave[2w],
// wave[2w+1] are the byte and value offsets of block 8w (one prefix per
// wavefront's group of eight blocks, from utl_launch_wave_index); each wave
// scans its eight meta words itself.
struct IndexRef {
  const uint64_t* boff;
  const uint64_t* ooff;
  const uint32_t* meta;
  const uint64_t* wave;
};

// Decodes the wave's blocks [b0, b0 + nb) (b0 a multiple of 8). The
// wave-uniform round logic stages as many consecutive blocks as fit the
// arena, decodes them, and repeats until every block is done.
template <bool U64, bool META, class WL>
__device__ __forceinline__ void decode_group(WL* W, const uint8_t* __restrict__ src, const IndexRef ix, uint64_t b0,
                                             int nb, void* __restrict__ out, int32_t* __restrict__ status,
                                             int32_t* __restrict__ consumed_out) {
  const int blk = lane_id() >> 3, j = lane_id() & 7;
  const bool mine = blk < nb;
  const uint64_t b = b0 + (mine ? blk : 0);
  // this lane's block: bytes [my_o0, my_o1) of src, slot_n values at out0
  uint64_t my_o0, my_o1, out0, slot_n;
  if (META) {
    // the group's wave-index entry (uniform: scalar loads) and the lanes'
    // meta words, both loads in flight at once (no branch between them: a
    // lane past the last block reads block b0's word and drops it)
    const uint64_t w = b0 >> 3;
    const uint64_t wb = ix.wave[2 * w], wv = ix.wave[2 * w + 1];
    const uint32_t m = ix.meta[b];
    const uint32_t ln = mine ? m & 0xFFFFFFu : 0u, cn = mine ? m >> 24 : 0u;
    const uint32_t il = wave_group_scan(ln), ic = wave_group_scan(cn);  // inclusive over the 8 blocks
    my_o0 = wb + (il - ln);
    my_o1 = my_o0 + ln;
    out0 = wv + (ic - cn);
    slot_n = cn;
  } else {
    my_o0 = ix.boff[b];
    my_o1 = ix.boff[b + 1];
    out0 = ix.ooff[b];
    slot_n = ix.ooff[b + 1] - out0;
  }
  const int64_t len = (int64_t)(my_o1 - my_o0);
  // Common case: the whole group fits the arena in one round: every block's
  // offsets are monotone (one ballot over the lanes' own offsets) and the
  // bytes from the first block's start to the last block's extent fit.
  const bool mono = __ballot(mine && my_o1 < my_o0) == 0;
  const int last_lane = kTpb * (nb - 1);
  const uint64_t o_first = readlane64(my_o0, 0);
  const uint64_t o_lastm1 = readlane64(my_o0, last_lane);
  const int64_t ext_all = clamp_extent<U64>((int64_t)(readlane64(my_o1, last_lane) - o_lastm1));
  constexpr int kArena = WL::kArenaBytes;
  const bool one_round = mono && (int64_t)(o_lastm1 - o_first) + ext_all + 15 + 4 <= kArena;
  int first = 0;
#pragma unroll 1
  while (first < nb) {
    uint64_t base;
    int last;
    int64_t span;  // staged bytes from base
    if (one_round) {
      base = o_first;
      last = nb;
      span = (int64_t)(o_lastm1 - o_first) + ext_all;
    } else {
      // Blocks from `first` on while each one's extent fits the arena. Their
      // starts need not increase (a gathered index jumps between requests),
      // so the round stages up to the furthest extent among them, not the
      // last one's.
      base = readlane64(my_o0, kTpb * first);
      last = first;
      span = 0;
#pragma unroll 1
      while (last < nb) {
        const uint64_t o0 = readlane64(my_o0, kTpb * last), o1 = readlane64(my_o1, kTpb * last);
        const int64_t ext = clamp_extent<U64>((int64_t)(o1 - o0));
        if (o0 < base || (int64_t)(o0 - base) + ext + 15 + 4 > kArena) break;
        span = max(span, (int64_t)(o0 - base) + ext);
        ++last;
      }
      if (last == first) {  // (block `first` always fits: its extent is at most kExtent64)
        last = first + 1;
        span = clamp_extent<U64>((int64_t)(readlane64(my_o1, kTpb * first) - base));
      }
    }
    const uint8_t* gbase = src + base;
    stage_range(W, gbase, span);
    int jj = j, bb = blk;  // opaque per round: keeps per-thread LDS addresses out of the round loop
    asm volatile("" : "+v"(jj), "+v"(bb));
    const bool active = mine && bb >= first && bb < last;
    const int at = (int)(reinterpret_cast<uintptr_t>(gbase) & 15) + (int)(my_o0 - base);
    int32_t st = UTL_OK;
    int consumed = 0;
    if (!U64) {
      uint32_t v0[kRows], v1[kRows];
      SubResult r1;
      st = decode_sub<false>(W, active, bb, active ? at : 0, len, jj, v0, v1, r1);
      bool put = false;
      if (active) {
        consumed = r1.consumed;
        if (st == UTL_OK && (uint64_t)r1.count > slot_n) st = UTL_ERR_ARGUMENT;
        put = st == UTL_OK && r1.count > 0;
        if (jj == 0) {
          status[b] = st;
          if (consumed_out) consumed_out[b] = st == UTL_OK ? consumed : 0;
        }
      }
      // Full blocks at 16-value offsets write whole segments by themselves;
      // a round with any other block goes through aux when its blocks fill
      // one range without holes. (Full blocks at other even offsets used to
      // store their 64-byte rows directly too, each row split across two
      // segments: on 200,000 uneven streams that wrote 8.7% more bytes to
      // HBM than the output holds, and staging them took the decode from
      // 4.35 to 4.11 ms.)
      const bool direct = direct_store(r1.count, out0);
      const bool holes = __ballot(active && !(st == UTL_OK && (uint64_t)r1.count == slot_n)) != 0;
      if (!holes && __ballot(put && !direct) != 0) {
        const uint64_t g0 = readlane64(out0, kTpb * first);
        const uint64_t g1 = readlane64(out0 + slot_n, kTpb * (last - 1));
        store_u32_staged(W, (uint32_t*)out, g0, g1, put, out0, r1.count, jj, v0, v1);
      } else if (put) {
        store_u32((uint32_t*)out, out0, r1.count, jj, v0, v1);
      }
      wave_fence();
    } else {
      // uint64.go:191-266. One decode_sub instance serves every pass: pass 0
      // decodes first sub-blocks, pass 1 the staged second sub-blocks of
      // two-block blocks, and each further pass one second sub-block that
      // lies outside the staged extent (corrupt blocks only), restaged.
      // decode_sub writes h only on the lanes that decode in a pass, so h
      // carries each block's second sub-block across the later passes.
      uint32_t l0[kRows], l1[kRows], h0[kRows], h1[kRows];
      uint64_t base64 = 0;
      Hdr h;
      bool go1;
      st = u64_prelude(W->arena, active, active ? at : 0, len, h, base64, go1);
      bool two = false, need2 = false, pending = false;
      int count = 0, c1 = 0, b2len = 0;
      {
        SubResult r;
        const int32_t s1 = decode_sub<true>(W, go1, bb, go1 ? at : 0, len, jj, l0, l1, r);
        if (go1) {
          st = s1;
          if (st == UTL_OK && !h.combine) {
            count = r.count;
            consumed = r.consumed;
          } else if (st == UTL_OK) {
            const bool he = h.exc > 0;
            const int forb = for_base_bytes(h.for_width);
            if (len < payload_offset(forb, he, true)) {
              st = UTL_ERR_INVALID_BUFFER;
            } else {
              b2len = (int)lds_bytes(W->arena, at + kHdr + (he ? kSvbLen : 0) + forb, 2);
              if ((int64_t)r.consumed + b2len > len) {
                st = UTL_ERR_INVALID_BUFFER;
              } else {
                two = true;
                count = r.count;
                c1 = r.consumed;
                if ((int64_t)c1 + clamp_extent<false>(b2len) > clamp_extent<true>(len))
                  pending = true;
                else
                  need2 = true;
              }
            }
          }
        }
      }
      if (__ballot(need2 || pending)) {
        // pass 1: the staged second sub-blocks; then one pass per second
        // sub-block outside the staged extent (corrupt blocks only), restaged
        bool go_p = need2;
        int at_p = need2 ? at + c1 : 0;
        uint64_t pmask = __ballot(active && st == UTL_OK && pending && jj == 0);
#pragma unroll 1
        for (;;) {
          int j2 = jj, b2 = bb;  // opaque per pass: keeps lane constants out of the loop
          asm volatile("" : "+v"(j2), "+v"(b2));
          SubResult r;
          const int32_t s2 = decode_sub<true>(W, go_p, b2, go_p ? at_p : 0, b2len, j2, h0, h1, r);
          wave_fence();
          if (go_p) {
            st = s2;
            if (st == UTL_OK && r.count != count) st = UTL_ERR_INVALID_BUFFER;
            if (st == UTL_OK) consumed = c1 + b2len;
          }
          if (pmask == 0) break;
          const int k = __ffsll((long long)pmask) - 1;  // lane of the block's thread 0
          pmask &= pmask - 1;
          const int ck = __shfl(c1, k, 64), bk = __shfl(b2len, k, 64);
          const uint8_t* g = src + readlane64(my_o0, k) + ck;
          stage_range(W, g, clamp_extent<false>(bk));
          go_p = bb == (k >> 3);
          at_p = (int)(reinterpret_cast<uintptr_t>(g) & 15);
        }
      }
      if (active) {
        if (st == UTL_OK && (uint64_t)count > slot_n) st = UTL_ERR_ARGUMENT;
        if (st == UTL_OK && count > 0)
          store_u64((uint64_t*)out, out0, count, jj, l0, l1, h0, h1, two, base64);
        if (jj == 0) {
          status[b] = st;
          if (consumed_out) consumed_out[b] = st == UTL_OK ? consumed : 0;
        }
      }
    }
    first = last;
  }
}

// Workgroup g runs on XCD g % 8; tile(g) makes each XCD's workgroups cover
// one contiguous range of tiles.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t g, uint64_t per_xcd) { return (g & 7) * per_xcd + (g >> 3); }

// Flat form: one workgroup per tile of 8 * WAVES blocks. OCC is the
// occupancy hint (waves per SIMD) the variant's LDS allows.
template <bool U64, bool META, int ARENA, int WAVES, int OCC>
__global__ __launch_bounds__(64 * WAVES, OCC) void k_decode(const uint8_t* __restrict__ src, const IndexRef ix,
                                                           uint64_t nblocks, void* __restrict__ out,
                                                           int32_t* __restrict__ status,
                                                           int32_t* __restrict__ consumed_out, uint64_t per_xcd) {
  __shared__ WaveLdsT<ARENA> lds[WAVES];
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint64_t tile = xcd_tile(blockIdx.x, per_xcd);
  const uint64_t b0 = tile * (kWaveBlocks * WAVES) + (uint64_t)wave * kWaveBlocks;
  if (b0 >= nblocks) return;
  const int nb = (int)min((uint64_t)kWaveBlocks, nblocks - b0);
  decode_group<U64, META>(&lds[wave], src, ix, b0, nb, out, status, consumed_out);
}

// Persistent form: a fixed grid; workgroups on XCD x take tiles from XCD x's
// own work counter (counter[8x]) over XCD x's contiguous range of tiles, so
// load balance stays inside the XCD. Every workgroup reaches the exit test.
template <bool U64, bool META, int ARENA, int WAVES, int OCC>
__global__ __launch_bounds__(64 * WAVES, OCC) void k_decode_persistent(
    const uint8_t* __restrict__ src, const IndexRef ix, uint64_t nblocks, void* __restrict__ out,
    int32_t* __restrict__ status, int32_t* __restrict__ consumed_out, unsigned long long* __restrict__ counter,
    uint64_t per_xcd, uint32_t tiles_per_grab) {
  __shared__ WaveLdsT<ARENA> lds[WAVES];
  __shared__ unsigned long long grab;
  constexpr uint64_t kTileBlocks = kWaveBlocks * WAVES;
  const uint32_t xcd = blockIdx.x & 7;
  const uint32_t wave = uniform(threadIdx.x >> 6);
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
      const uint64_t b0 = tile * kTileBlocks + (uint64_t)wave * kWaveBlocks;
      if (b0 < nblocks) {
        const int nb = (int)min((uint64_t)kWaveBlocks, nblocks - b0);
        decode_group<U64, META>(&lds[wave], src, ix, b0, nb, out, status, consumed_out);
      }
    }
  }
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// One launch of variant (U64, ARENA, WAVES) in either index form: the flat
// grid, or the persistent grid of grid_wgs workgroups when counter is not
// null.
template <bool U64, int ARENA, int WAVES, int OCC, int OCC_P>
hipError_t launch_variant(const uint8_t* src, const IndexRef& ix, uint64_t nblocks, void* out, int32_t* status,
                          int32_t* consumed, unsigned long long* counter, int grid_wgs, uint32_t grab,
                          hipStream_t stream) {
  const uint64_t ntiles = (nblocks + kWaveBlocks * WAVES - 1) / (kWaveBlocks * WAVES);
  const uint64_t per_xcd = (ntiles + 7) / 8;
  const dim3 block(64 * WAVES);
  const bool meta = ix.boff == nullptr;
  if (!counter) {
    const dim3 grid((unsigned)(per_xcd * 8));
    if (meta)
      k_decode<U64, true, ARENA, WAVES, OCC><<<grid, block, 0, stream>>>(src, ix, nblocks, out, status, consumed,
                                                                         per_xcd);
    else
      k_decode<U64, false, ARENA, WAVES, OCC><<<grid, block, 0, stream>>>(src, ix, nblocks, out, status, consumed,
                                                                          per_xcd);
  } else {
    k_fill_u64<<<1, 64, 0, stream>>>(counter, 0ull, 64);
    const dim3 grid((unsigned)((grid_wgs + 7) / 8 * 8));
    grab = grab ? grab : 1;
    if (meta)
      k_decode_persistent<U64, true, ARENA, WAVES, OCC_P><<<grid, block, 0, stream>>>(
          src, ix, nblocks, out, status, consumed, counter, per_xcd, grab);
    else
      k_decode_persistent<U64, false, ARENA, WAVES, OCC_P><<<grid, block, 0, stream>>>(
          src, ix, nblocks, out, status, consumed, counter, per_xcd, grab);
  }
  return hipGetLastError();
}

}  // namespace utl

using namespace utl;

bool utl_tune_valid(const utl_tune& t) {
  const bool arena = t.arena_bytes == kArena32Small || t.arena_bytes == kArena32Mid || t.arena_bytes == kArena32Large;
  return arena && (t.waves_per_wg == 2 || t.waves_per_wg == 4) && t.grid_wgs >= 1 && t.tiles_per_grab >= 1;
}

// The occupancy hints follow the LDS each variant takes per CU (160 KiB):
// 2560 B: 24 waves (6 per SIMD; the persistent form holds 5 without spills),
// 3328 B: 20-22 (5), 4352 B: 16-18 (4); u64: 12 waves, hinted 4.
static hipError_t launch_any(bool u64, const uint8_t* src, const IndexRef& ix, uint64_t nblocks, void* out,
                             int32_t* status, int32_t* consumed, unsigned long long* counter, const utl_tune& t,
                             hipStream_t stream) {
  if (nblocks == 0) return hipSuccess;
  const int g = t.grid_wgs;
  const uint32_t q = t.tiles_per_grab;
  if (u64) {
    if (t.waves_per_wg == 2)
      return launch_variant<true, kArena64, 2, 4, 4>(src, ix, nblocks, out, status, consumed, counter, g, q, stream);
    return launch_variant<true, kArena64, 4, 4, 4>(src, ix, nblocks, out, status, consumed, counter, g, q, stream);
  }
  switch (t.arena_bytes * 8 + t.waves_per_wg) {
    case kArena32Mid * 8 + 2:
      return launch_variant<false, kArena32Mid, 2, 5, 5>(src, ix, nblocks, out, status, consumed, counter, g, q,
                                                         stream);
    case kArena32Mid * 8 + 4:
      return launch_variant<false, kArena32Mid, 4, 5, 5>(src, ix, nblocks, out, status, consumed, counter, g, q,
                                                         stream);
    case kArena32Large * 8 + 2:
      return launch_variant<false, kArena32Large, 2, 4, 4>(src, ix, nblocks, out, status, consumed, counter, g, q,
                                                           stream);
    case kArena32Large * 8 + 4:
      return launch_variant<false, kArena32Large, 4, 4, 4>(src, ix, nblocks, out, status, consumed, counter, g, q,
                                                           stream);
    case kArena32Small * 8 + 2:
      return launch_variant<false, kArena32Small, 2, 6, 5>(src, ix, nblocks, out, status, consumed, counter, g, q,
                                                           stream);
    default:
      return launch_variant<false, kArena32Small, 4, 6, 5>(src, ix, nblocks, out, status, consumed, counter, g, q,
                                                           stream);
  }
}

hipError_t utl_launch_decode(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                             void* out, int32_t* status, int32_t* consumed, const utl_tune& tune, hipStream_t stream) {
  const IndexRef ix{boff, ooff, nullptr, nullptr};
  return launch_any(u64, src, ix, nblocks, out, status, consumed, nullptr, tune, stream);
}

hipError_t utl_launch_decode_persistent(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff,
                                        uint64_t nblocks, void* out, int32_t* status, int32_t* consumed,
                                        unsigned long long* counter, const utl_tune& tune, hipStream_t stream) {
  const IndexRef ix{boff, ooff, nullptr, nullptr};
  return launch_any(u64, src, ix, nblocks, out, status, consumed, counter, tune, stream);
}

hipError_t utl_launch_decode_meta(bool u64, const uint8_t* src, const uint32_t* meta, const uint64_t* wave,
                                  uint64_t nblocks, void* out, int32_t* status, int32_t* consumed,
                                  unsigned long long* counter, const utl_tune& tune, hipStream_t stream) {
  const IndexRef ix{nullptr, nullptr, meta, wave};
  return launch_any(u64, src, ix, nblocks, out, status, consumed, counter, tune, stream);
}
