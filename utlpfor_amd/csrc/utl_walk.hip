// utl_walk.hip -- index of foreign streams whose block starts are unknown.
//
// The reference finds blocks by a serial walk: block k+1 starts
// BlockLength(block k) bytes after block k (block_length.go:8,
// example_test.go:171-187). The walk of stream s stops at the stream's end
// (status 0), at the first BlockLength error (that error) or at a block
// longer than the bytes left (ErrInvalidBuffer). The kernels here give the
// same blocks and statuses.
//
// Short streams (< kLong bytes): one lane per stream walks it, reading the
// headers from global memory (k_walk_serial); a short stream whose first
// block is short enough that its walk would be a long chain of dependent
// loads (more than kDenseChain blocks) is walked by a workgroup, window by
// window, by pointer jumping instead (k_walk_dense).
//
// Long streams are cut into chunks and walked chunk-parallel. A stream's
// chunk length is the call's (walk_chunk_len bytes), shortened for a
// stream whose first block is short, so that a chunk holds about
// kChainBlocks blocks, but never below kChunk bytes (k_walk_plan; then
// k_walk_owner maps each chunk to its stream):
//   k_walk_vote    per chunk, a guess of its entry (the first block start at
//                  or after the chunk start): every candidate position of a
//                  kWarm-byte warm-up region before the chunk is followed by
//                  pointer jumping in LDS to the point where it crosses into
//                  the chunk, and the crossing most candidates reach wins
//                  (walks from wrong positions die at a failed header or
//                  merge into the stream's own chain);
//   k_walk_chunks  one lane per chunk walks it from the guess, reading the
//                  headers from global memory: block count, exit (first
//                  position at or past the chunk end) and stop status; the
//                  chunks of streams of tiny blocks go to k_walk_chunks_pj
//                  (a workgroup per chunk, pointer jumping);
//   k_walk_verify  one workgroup per long stream (64 threads for streams of
//                  at most 64 chunks, else 1024) checks, a workgroup's
//                  width of chunks at a time, that each chunk's guess
//                  equals the previous chunk's exit;
//                  the first chunk that does not is walked again from that
//                  exit. The result is exact whatever the guesses were: a
//                  chunk counts only once its entry is the true one, and the
//                  first chunk's entry is the stream start.
//   k_walk_emit    (second pass) one lane per live chunk walks it again and
//                  writes its blocks' meta words and starts (k_walk_chunks_pj
//                  for the chunks of streams of tiny blocks).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "utl_device.h"
#include "utl_format.h"
#include "utl_kernels.h"

namespace utl {
namespace {

constexpr int kChunk = 16320;            // least bytes of a long stream per chunk
constexpr int kWarm = 2048;              // warm-up bytes before a chunk
constexpr int kLand = 2048;              // entry candidates (the warm-up's first bytes); above every block
                                         // length an encoder emits (2038)
constexpr uint64_t kLong = 2 * kChunk;   // streams at least this long are walked chunk-parallel
constexpr int kWin16 = 256;              // LDS window of a walking wavefront (4 KiB), in 16-byte units
constexpr int kStep = kWin16 * 16 - 64;  // bytes walked per window (the rest: header lookahead, alignment)
constexpr int kWarm16 = (kWarm + 64) / 16;  // LDS window of one vote
constexpr uint64_t kNone = ~0ull;
constexpr uint16_t kDead = 0xFFFF;
constexpr uint32_t kStLand = 0, kStEnd = 1, kStStop = 0x100;  // chunk exit kinds (kStStop | error)
constexpr int kSmallVer = 64;  // long streams of at most this many chunks are verified by 64 threads
constexpr uint64_t kChainBlocks = 280;  // blocks a chunk is sized for, from the stream's first block
constexpr uint64_t kDenseChain = 320;   // short streams of more blocks than this (estimated) walk by wavefront
constexpr uint32_t kDense = 1;          // clen of a short stream walked by a workgroup (else 0)
constexpr uint32_t kClenPJ = 0x80000000u;  // clen flag: the stream's chunks are walked by workgroups

__device__ __forceinline__ uint64_t chunk_len_of(uint32_t cl) { return cl & ~kClenPJ; }

struct WalkRec {
  uint64_t e;     // entry (stream-relative), kNone if the vote found none
  uint64_t x;     // exit: first chain position at or past the chunk end (kStLand)
  uint32_t n;     // blocks starting in [e, x)
  uint32_t st;    // kStLand, kStEnd or kStStop | error
  uint32_t base;  // blocks of the stream before this chunk (written by k_walk_verify)
  uint32_t pad;
};
static_assert(sizeof(WalkRec) == 32, "WalkRec layout");
static_assert(kLand <= kWarm, "the vote reads kLand jump entries");

// BlockLength (block_length.go:8) of the block whose first bytes are
// h = bytes 0..3, d1 = bytes 4..7, d2 = bytes 8..11 (little-endian), with
// `left` bytes from the block start to the end of its stream. Branch-free;
// the error precedence is the reference's: short header, count > 128,
// bit width > 32, invalid flags, short svbLen field, short block2Len field.
__device__ __forceinline__ int step_len(uint32_t h, uint32_t d1, uint32_t d2, int64_t left, uint32_t* bl) {
  const uint32_t t = (h >> 13) & 3u, fw = (h >> 15) & 3u, exc = h >> 24, bwf = (h >> 8) & 31u;
  const bool t3 = t == 3u;
  const bool combine = t3 && (h & kHCombine) != 0u;
  const uint32_t bad = h & (kHReserved | kHBlock256 | (t3 ? 0u : kHCombine));
  const uint32_t forb = (t3 && !combine && fw == 3u) ? (uint32_t)kFor64 : ((1u << fw) >> 1);
  const bool he = exc != 0u;
  const uint32_t b2off = (uint32_t)kHdr + (he ? (uint32_t)kSvbLen : 0u) + forb;  // block2Len (combine: <= 10)
  const uint32_t poff = b2off + (combine ? (uint32_t)kB2Len : 0u);
  const uint32_t idx = exc <= (uint32_t)kExcBitmap ? exc : (uint32_t)kExcBitmap;
  uint32_t b = poff + (bwf << 6) + (he ? idx + (d1 & 0xFFFFu) : 0u);
  const uint64_t q = ((uint64_t)d2 << 32) | d1;  // bytes 4..11
  const uint32_t sh = combine ? 8u * (b2off - 4u) : 0u;
  b += combine ? ((uint32_t)(q >> sh) & 0xFFFFu) : 0u;
  *bl = b;
  int err = UTL_OK;
  if (combine && left < (int64_t)poff) err = UTL_ERR_INVALID_BUFFER;
  if (he && left < kHdr + kSvbLen) err = UTL_ERR_INVALID_BUFFER;
  if (bad) err = UTL_ERR_INVALID_FLAGS;
  if (bwf > 8u) err = UTL_ERR_INVALID_BUFFER;
  if ((h & 0xFFu) > (uint32_t)kBlock) err = UTL_ERR_INVALID_BLOCK_LENGTH;
  if (left < kHdr) err = UTL_ERR_INVALID_BUFFER;
  return err;
}

// Stages the bytes [lo, hi) of src into W (CAP16 16-byte units) from the
// 16-byte boundary at or below lo; no byte at or past hi is read (those
// slots read as zero). Returns the window's absolute base. Threads
// [0, NTHR) of the caller take part, each with all its loads in flight at
// once; the caller synchronizes afterwards.
template <int CAP16, int NTHR>
__device__ __forceinline__ uint64_t stage(uint4* W, const uint8_t* __restrict__ src, uint64_t lo, uint64_t hi,
                                          int tid) {
  constexpr int U = (CAP16 + NTHR - 1) / NTHR;
  const uint64_t base = lo & ~15ull;
  const uint64_t span = hi > base ? hi - base : 0;
  const int nfull = (int)min(span >> 4, (uint64_t)CAP16);  // whole 16-byte units below hi
  uint4 v[U];
  if (nfull > 0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = min(u * NTHR + tid, nfull - 1);
      v[u] = *reinterpret_cast<const uint4*>(src + base + 16ull * (uint64_t)c);
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int c = u * NTHR + tid;
    if (c < CAP16 && c != nfull) W[c] = c < nfull ? v[u] : make_uint4(0u, 0u, 0u, 0u);
  }
  static_assert(NTHR >= 16, "the unit holding hi is read a byte per thread");
  if (tid < 16 && nfull < CAP16) {  // the unit holding hi: its bytes below hi, zeros after
    const uint64_t b = base + 16ull * (uint64_t)nfull + (uint64_t)tid;
    reinterpret_cast<uint8_t*>(W + nfull)[tid] = b < hi ? src[b] : (uint8_t)0;
  }
  return base;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bytes 0..11 from window byte i as three dwords (one batch of LDS reads).
__device__ __forceinline__ void win96(const uint32_t* W32, int i, uint32_t& h, uint32_t& d1, uint32_t& d2) {
  const int w = i >> 2;
  const uint32_t s = (uint32_t)(i & 3);
  const uint32_t a0 = W32[w], a1 = W32[w + 1], a2 = W32[w + 2], a3 = W32[w + 3];
  h = __builtin_amdgcn_alignbyte(a1, a0, s);
  d1 = __builtin_amdgcn_alignbyte(a2, a1, s);
  d2 = __builtin_amdgcn_alignbyte(a3, a2, s);
}

struct WalkRes {
  uint64_t x;
  uint32_t n, st;
};

// BlockLength of a block at least 16 bytes from its stream's end, where
// none of the length checks can fail: the length, and whether the header
// passes the count, bit-width and flag checks.
__device__ __forceinline__ uint32_t fast_len(uint32_t h, uint32_t d1, uint32_t d2, bool* ok) {
  const uint32_t t = (h >> 13) & 3u, fw = (h >> 15) & 3u, exc = h >> 24, bwf = (h >> 8) & 31u;
  const bool t3 = t == 3u;
  const bool combine = t3 && (h & kHCombine) != 0u;
  const uint32_t bad = h & (kHReserved | kHBlock256 | (t3 ? 0u : kHCombine));
  const uint32_t forb = (t3 && !combine && fw == 3u) ? (uint32_t)kFor64 : ((1u << fw) >> 1);
  const bool he = exc != 0u;
  const uint32_t b2off = (uint32_t)kHdr + (he ? (uint32_t)kSvbLen : 0u) + forb;
  uint32_t b = b2off + (bwf << 6) + (he ? min(exc, (uint32_t)kExcBitmap) + (d1 & 0xFFFFu) : 0u);
  if (combine) b += (uint32_t)kB2Len + ((uint32_t)((((uint64_t)d2 << 32) | d1) >> (8u * (b2off - 4u))) & 0xFFFFu);
  *ok = (h & 0xFFu) <= (uint32_t)kBlock && bwf <= 8u && bad == 0u;
  return b;
}

// Walks a stream (absolute start so, length Ls) from position pos while
// pos < cend (cend - pos <= kStep). The window W32 (absolute base wabs)
// holds the stream bytes [pos, cend + 16) wherever they exist. With meta
// set, lane 0 writes block i's meta word to meta[out + i] and its absolute
// start to start[out + i]. Offsets within the window are 32-bit.
__device__ WalkRes walk_range(const uint32_t* W32, uint64_t wabs, uint64_t so, uint64_t Ls, uint64_t pos,
                              uint64_t cend, uint32_t* __restrict__ meta, uint64_t* __restrict__ start,
                              uint64_t out) {
  // bytes from pos to the stream end, capped: no block is 2^31 bytes long,
  // so a capped value never ends the stream or fails a length check
  const uint32_t lim = (uint32_t)min(Ls - pos, (uint64_t)0x7FFFFFFF);
  const uint32_t wlen = (uint32_t)(cend - pos);
  const int i0 = (int)(so + pos - wabs);
  uint32_t n = 0, st = kStLand, off = 0;
  while (off < wlen) {
    uint32_t h, d1, d2;
    win96(W32, i0 + (int)off, h, d1, d2);
    const uint32_t left = lim - off;
    bool ok;
    uint32_t bl = fast_len(h, d1, d2, &ok);
    if (!ok || left < 16u) {  // the reference's checks in order
      const int err = step_len(h, d1, d2, (int64_t)left, &bl);
      if (err != UTL_OK) {
        st = kStStop | (uint32_t)err;
        break;
      }
    }
    if (bl > left) {
      st = kStStop | UTL_ERR_INVALID_BUFFER;
      break;
    }
    if (meta && lane_id() == 0) {
      meta[out + n] = (bl & 0xFFFFFFu) | ((h & 0xFFu) << 24);
      if (start) start[out + n] = so + pos + off;
    }
    ++n;
    off += bl;
    if (off == lim) {
      st = kStEnd;
      break;
    }
  }
  return {pos + off, n, st};
}

// Walks from pos while pos < end through kStep-byte LDS windows (a
// wave-collective call; the verification's re-walk of a chunk whose entry
// vote missed).
__device__ WalkRes walk_span(uint4* W, const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls, uint64_t pos,
                             uint64_t end, uint32_t* __restrict__ meta, uint64_t* __restrict__ start, uint64_t out) {
  uint32_t n = 0;
  while (pos < end) {
    const uint64_t wend = min(pos + (uint64_t)kStep, end);
    const uint64_t wabs = stage<kWin16, 64>(W, src, so + pos, so + min(wend + 32, Ls), lane_id());
    wave_sync();
    const WalkRes r =
        walk_range(reinterpret_cast<const uint32_t*>(W), wabs, so, Ls, pos, wend, meta, start, out + n);
    wave_sync();  // W is restaged next
    n += r.n;
    pos = r.x;
    if (r.st != kStLand) return {pos, n, r.st};
  }
  return {pos, n, kStLand};
}

// Walks chunk k of a long stream from pos (a wave-collective call).
__device__ __forceinline__ WalkRes walk_chunk(uint4* W, const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls,
                                              uint64_t chunk, uint64_t k, uint64_t pos, uint32_t* __restrict__ meta,
                                              uint64_t* __restrict__ start, uint64_t out) {
  return walk_span(W, src, so, Ls, pos, min((k + 1) * chunk, Ls), meta, start, out);
}

// The dwords holding bytes 0..11 of a stream from stream byte pos, read
// from global memory by one lane: the four aligned dwords around them when
// the stream has at least 16 bytes from pos on (sh: the byte offset in the
// first), else the bytes up to the stream's end with zeros after (sh 0), so
// no byte past the end is read. src is 16-byte aligned, so the aligned
// dwords never start before it. The words are only combined by hdr96, which
// lets a caller issue the load and use its registers later.
struct Raw96 {
  uint32_t a0, a1, a2, a3, sh;
};

__device__ __forceinline__ Raw96 load96(const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls, uint64_t pos) {
  const uint64_t a = so + pos;
  Raw96 r;
  if (Ls - pos >= 16) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(src + (a & ~3ull));
    r.a0 = w[0];
    r.a1 = w[1];
    r.a2 = w[2];
    r.a3 = w[3];
    r.sh = (uint32_t)(a & 3);
  } else {
    const int n = (int)(Ls - pos);  // 1..15
    uint32_t b[3] = {0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < 12; ++k)
      if (k < n) b[k >> 2] |= (uint32_t)src[a + k] << (8 * (k & 3));
    r.a0 = b[0];
    r.a1 = b[1];
    r.a2 = b[2];
    r.a3 = 0u;
    r.sh = 0u;
  }
  return r;
}

__device__ __forceinline__ void hdr96(const Raw96& r, uint32_t& h, uint32_t& d1, uint32_t& d2) {
  h = __builtin_amdgcn_alignbyte(r.a1, r.a0, r.sh);
  d1 = __builtin_amdgcn_alignbyte(r.a2, r.a1, r.sh);
  d2 = __builtin_amdgcn_alignbyte(r.a3, r.a2, r.sh);
}

// What a lane walk writes per block: nothing (counts only), the meta word,
// or the meta word and the block's absolute start.
enum LaneOut { kCount = 0, kMeta = 1, kMetaStart = 2 };

// A lane walk's last outputs, stored four blocks at a time: global loads
// and stores share one in-order counter, so a store issued before the next
// header's load delays it; with the stores batched, three steps in four
// wait for their load alone. q0 is the newest entry.
struct OutQ {
  uint32_t m0, m1, m2, m3;
  uint64_t s0, s1, s2, s3;
  uint32_t k;  // entries held (0..3 between steps)
};

// Stores the k newest entries of q, the newest at index last.
template <int OUT>
__device__ __forceinline__ void outq_flush(const OutQ& q, uint32_t k, uint32_t* __restrict__ meta,
                                           uint64_t* __restrict__ start, uint64_t last) {
  if (OUT >= kMeta) {
    if (k > 0) meta[last] = q.m0;
    if (k > 1) meta[last - 1] = q.m1;
    if (k > 2) meta[last - 2] = q.m2;
    if (k > 3) meta[last - 3] = q.m3;
  }
  if (OUT >= kMetaStart) {
    if (k > 0) start[last] = q.s0;
    if (k > 1) start[last - 1] = q.s1;
    if (k > 2) start[last - 2] = q.s2;
    if (k > 3) start[last - 3] = q.s3;
  }
}

// One block of a lane walk: the header in r (at pos) is checked, the next
// header's load is issued, then the block's outputs are queued (and the
// queue stored when it holds four). Returns whether the walk goes on (st
// is set when it stops).
template <int OUT>
__device__ __forceinline__ bool lane_step(const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls, uint64_t end,
                                          uint32_t* __restrict__ meta, uint64_t* __restrict__ start, uint64_t out,
                                          Raw96& r, uint64_t& pos, uint32_t& n, uint32_t& st, OutQ& q) {
  uint32_t h, d1, d2;
  hdr96(r, h, d1, d2);
  const uint64_t left = Ls - pos;
  bool ok;
  uint32_t bl = fast_len(h, d1, d2, &ok);
  if (!ok || left < 16u) {  // the reference's checks in order
    const int err = step_len(h, d1, d2, (int64_t)left, &bl);
    if (err != UTL_OK) {
      st = kStStop | (uint32_t)err;
      return false;
    }
  }
  if (bl > left) {
    st = kStStop | UTL_ERR_INVALID_BUFFER;
    return false;
  }
  const uint64_t next = pos + bl;
  const bool more = next < end && next != Ls;
  if (more) r = load96(src, so, Ls, next);
  if (OUT >= kMeta) {
    q.m3 = q.m2;
    q.m2 = q.m1;
    q.m1 = q.m0;
    q.m0 = (bl & 0xFFFFFFu) | ((h & 0xFFu) << 24);
    if (OUT >= kMetaStart) {
      q.s3 = q.s2;
      q.s2 = q.s1;
      q.s1 = q.s0;
      q.s0 = so + pos;
    }
    if (++q.k == 4) {
      outq_flush<OUT>(q, 4, meta, start, out + n);
      q.k = 0;
    }
  }
  ++n;
  pos = next;
  if (!more && pos == Ls) st = kStEnd;
  return more;
}

// walk_span for one lane: walks the stream (absolute start so, length Ls)
// from pos while pos < end, reading each header from global memory. OUT
// kMeta writes block i's meta word to meta[out + i], kMetaStart also its
// absolute start to start[out + i]. Every step advances by at least 4
// bytes. Each step issues the next header's load before its own stores and
// combines the words after them: one counter tracks loads and stores in
// issue order, so the wait for the load leaves the stores in flight. The
// first step is peeled so that every entry to the loop has that same order.
template <int OUT>
__device__ WalkRes walk_lane(const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls, uint64_t pos, uint64_t end,
                             uint32_t* __restrict__ meta, uint64_t* __restrict__ start, uint64_t out) {
  uint32_t n = 0, st = kStLand;
  if (pos >= end) return {pos, n, st};
  OutQ q;
  q.m0 = q.m1 = q.m2 = q.m3 = 0u;
  q.s0 = q.s1 = q.s2 = q.s3 = 0u;
  q.k = 0u;
  Raw96 r = load96(src, so, Ls, pos);
  if (lane_step<OUT>(src, so, Ls, end, meta, start, out, r, pos, n, st, q))
    while (lane_step<OUT>(src, so, Ls, end, meta, start, out, r, pos, n, st, q)) {
    }
  if (OUT >= kMeta && q.k) outq_flush<OUT>(q, q.k, meta, start, out + n - 1);
  return {pos, n, st};
}

// walk_lane with the output form chosen by which pointers are set.
__device__ __forceinline__ WalkRes walk_lane_out(const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls,
                                                 uint64_t pos, uint64_t end, uint32_t* __restrict__ meta,
                                                 uint64_t* __restrict__ start, uint64_t out) {
  if (!meta) return walk_lane<kCount>(src, so, Ls, pos, end, nullptr, nullptr, 0);
  if (!start) return walk_lane<kMeta>(src, so, Ls, pos, end, meta, nullptr, out);
  return walk_lane<kMetaStart>(src, so, Ls, pos, end, meta, start, out);
}

__device__ __forceinline__ uint64_t stream_len(const uint64_t* stream_off, uint64_t s, uint64_t* so) {
  const uint64_t a = stream_off[s], b = stream_off[s + 1];
  *so = a;
  return b > a ? b - a : 0;
}


// A stream is walked chunk-parallel when it is long and its chunks fit the
// record scratch (max_chunks); every kernel applies this same test.
__device__ __forceinline__ bool is_long(uint64_t Ls, const uint64_t* chunk_base, uint64_t s, uint64_t max_chunks) {
  return Ls >= kLong && chunk_base[s + 1] <= max_chunks;
}

__device__ __forceinline__ int32_t status_of(uint32_t st) {
  return st >= kStStop ? (int32_t)(st & 0xFFu) : UTL_OK;
}

// Chunk offsets of the streams: chunks per stream (0 for short streams),
// scanned within tiles of kPlan streams by k_walk_plan; with more than one
// tile, k_walk_scan scans the tile totals and k_walk_apply adds them.
constexpr int kPlan = 1024;

// chunk_base[s] = chunks of the streams of s's tile before s (s <= ns; the
// entry ns closes the last stream); tile_sum[t] = tile t's total; clen[s] =
// the chunk length of a long stream, kDense for a short stream walked by a
// workgroup, 0 for the others; flen[s] = a long stream's first block length
// (0 when unknown). The lengths come from each stream's first
// block (BlockLength where its header passes the checks that do not depend
// on the bytes left; no estimate otherwise): only the chunking depends on
// them, never the result.
__global__ __launch_bounds__(kPlan) void k_walk_plan(const uint8_t* __restrict__ src,
                                                     const uint64_t* __restrict__ stream_off, uint64_t ns,
                                                     uint64_t chunk, uint64_t* __restrict__ chunk_base,
                                                     uint64_t* __restrict__ tile_sum, uint32_t* __restrict__ clen,
                                                     uint32_t* __restrict__ flen, uint32_t* __restrict__ list,
                                                     uint32_t* __restrict__ dlist, uint32_t* __restrict__ cnt) {
  __shared__ uint64_t red[kPlan / 64];
  const int tid = threadIdx.x, wid = tid >> 6;
  const uint64_t s = (uint64_t)blockIdx.x * kPlan + (uint64_t)tid;
  uint64_t so;
  const uint64_t Ls = s < ns ? stream_len(stream_off, s, &so) : 0;
  uint64_t bl0 = 0;  // the first block's length, 0 when unknown
  if (Ls >= 16) {
    uint32_t h, d1, d2;
    hdr96(load96(src, so, Ls, 0), h, d1, d2);
    bool ok;
    const uint32_t bl = fast_len(h, d1, d2, &ok);
    bl0 = ok ? bl : 0;
  }
  uint64_t cl = chunk;
  if (Ls >= kLong && bl0) cl = min(chunk, max((uint64_t)kChunk, (kChainBlocks * bl0) & ~15ull));
  const uint64_t c = Ls >= kLong ? (Ls + cl - 1) / cl : 0;
  const bool dense = Ls < kLong && bl0 && Ls > kDenseChain * bl0;
  // blocks so small that even a kChunk-byte chunk is a long chain: the
  // chunks are walked by workgroups (pointer jumping) instead of lanes
  const bool pj = c && bl0 && kChainBlocks * bl0 < (uint64_t)kChunk;
  if (s < ns) {
    clen[s] = c ? (uint32_t)cl | (pj ? kClenPJ : 0u) : dense ? kDense : 0u;
    flen[s] = c ? (uint32_t)bl0 : 0u;
  }
  const uint64_t incl = wave_incl_scan64(c);
  if (lane_id() == 63) red[wid] = incl;
  __syncthreads();
  uint64_t before = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < kPlan / 64; ++w) {
    before += w < wid ? red[w] : 0u;
    sum += red[w];
  }
  if (s <= ns) chunk_base[s] = before + incl - c;
  if (tid == 0) tile_sum[blockIdx.x] = sum;
  // the long streams by verification class: list[0 ..) those of at most
  // kSmallVer chunks, list[ns - 1] downwards the others; the dense short
  // streams in dlist (one atomic per wavefront and class)
  const int ln = lane_id();
  const uint64_t lt = (1ull << ln) - 1;
  const uint64_t bs = __ballot(c != 0 && c <= (uint64_t)kSmallVer), bb = __ballot(c > (uint64_t)kSmallVer);
  const uint64_t bd = __ballot(dense);
  uint32_t a = 0, b = 0, e = 0;
  if (ln == 0 && bs) a = atomicAdd(&cnt[0], (uint32_t)__popcll(bs));
  if (ln == 0 && bb) b = atomicAdd(&cnt[1], (uint32_t)__popcll(bb));
  if (ln == 0 && bd) e = atomicAdd(&cnt[2], (uint32_t)__popcll(bd));
  a = (uint32_t)__builtin_amdgcn_readfirstlane((int)a);
  b = (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
  e = (uint32_t)__builtin_amdgcn_readfirstlane((int)e);
  if ((bs >> ln) & 1) list[a + (uint32_t)__popcll(bs & lt)] = (uint32_t)s;
  if ((bb >> ln) & 1) list[ns - 1 - (b + (uint32_t)__popcll(bb & lt))] = (uint32_t)s;
  if ((bd >> ln) & 1) dlist[e + (uint32_t)__popcll(bd & lt)] = (uint32_t)s;
}

__global__ void k_walk_zero(uint32_t* cnt) { cnt[threadIdx.x] = 0u; }  // the four list counters

// Exclusive scan of q[0..n) in place (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void k_walk_scan(uint64_t* q, uint64_t n) {
  __shared__ uint64_t part[1024];
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t a = min((uint64_t)threadIdx.x * per, n), b = min(a + per, n);
  uint64_t sum = 0;
  for (uint64_t i = a; i < b; ++i) sum += q[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const uint64_t v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = part[threadIdx.x] - sum;
  for (uint64_t i = a; i < b; ++i) {
    const uint64_t v = q[i];
    q[i] = run;
    run += v;
  }
}

// Adds the offsets of tiles 1.. to their streams' entries.
__global__ __launch_bounds__(kPlan) void k_walk_apply(uint64_t* __restrict__ chunk_base, uint64_t ns,
                                                      const uint64_t* __restrict__ tile_off) {
  const uint64_t t = (uint64_t)blockIdx.x + 1;
  const uint64_t s = t * kPlan + threadIdx.x;
  if (s <= ns) chunk_base[s] += tile_off[t];
}

// owner_of[c] = the stream of chunk c, for every chunk of the long streams
// (chunk_base final). A thread per stream fills a stream of at most 64
// chunks; the wavefront fills the longer ones of its 64 streams together.
__global__ __launch_bounds__(kPlan) void k_walk_owner(const uint64_t* __restrict__ chunk_base, uint64_t ns,
                                                      uint64_t max_chunks, uint32_t* __restrict__ owner_of) {
  const uint64_t s = (uint64_t)blockIdx.x * kPlan + threadIdx.x;
  uint64_t a = 0, b = 0;
  if (s < ns) {
    a = min(chunk_base[s], max_chunks);
    b = min(chunk_base[s + 1], max_chunks);
  }
  const bool big = b - a > 64;
  if (!big)
    for (uint64_t c = a; c < b; ++c) owner_of[c] = (uint32_t)s;
  for (uint64_t m = __ballot(big); m; m &= m - 1) {  // (all 64 lanes)
    const int l = __builtin_ctzll(m);
    const uint64_t al = __shfl(a, l, 64), bl = __shfl(b, l, 64), sl = __shfl(s, l, 64);
    for (uint64_t c = al + (uint64_t)lane_id(); c < bl; c += 64) owner_of[c] = (uint32_t)sl;
  }
}

// Whether header h passes the checks of BlockLength that do not depend on
// the bytes left (count, bit width, flags). A header failing them ends
// every walk that reaches it.
__device__ __forceinline__ bool hdr_ok(uint32_t h) {
  const bool t3 = ((h >> 13) & 3u) == 3u;
  const uint32_t bad = h & (kHReserved | kHBlock256 | (t3 ? 0u : kHCombine));
  return (h & 0xFFu) <= (uint32_t)kBlock && ((h >> 8) & 31u) <= 8u && bad == 0u;
}

constexpr int kVotePer = kWarm / 256;  // warm-up positions per thread of a vote
static_assert(kVotePer == 8 && kLand == kWarm, "a vote thread owns 8 consecutive positions, one 16-byte J row");

// Entry vote of chunk k >= 1 (one workgroup of 256 threads per chunk;
// thread t owns the warm-up positions 8t .. 8t+7). The header checks run on
// all positions from three shared dwords; BlockLength is computed only for
// the positions that pass them, and pointer jumping and voting loop over a
// thread's live positions only (a few per cent of all positions on packed
// data: every other position is dead after one hop).
__global__ __launch_bounds__(256) void k_walk_vote(const uint8_t* __restrict__ src,
                                                   const uint64_t* __restrict__ stream_off, uint64_t ns,
                                                   const uint64_t* __restrict__ chunk_base, uint64_t max_chunks,
                                                   const uint32_t* __restrict__ owner_of,
                                                   const uint32_t* __restrict__ clen, const uint32_t* __restrict__ flen,
                                                   WalkRec* __restrict__ recs) {
  __shared__ uint64_t fast_e;
  __shared__ uint4 W[kWarm16];
  __shared__ uint4 J4[kWarm / 8];  // J[p]: next hop of position p (16-bit entries)
  __shared__ uint32_t hist[kLand];
  __shared__ uint32_t best[4];
  uint16_t* J = reinterpret_cast<uint16_t*>(J4);
  const uint64_t total = min(chunk_base[ns], max_chunks);
  const int tid = threadIdx.x;
  const int p0 = kVotePer * tid;
  // cleared once; each chunk clears the entries it counted in
  for (int i = tid; i < kLand; i += 256) hist[i] = 0;
  for (uint64_t c = blockIdx.x; c < total; c += gridDim.x) {
    const uint64_t s = owner_of[c];
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    if (!is_long(Ls, chunk_base, s, max_chunks)) continue;
    const uint64_t k = c - chunk_base[s];
    if (k == 0) {
      if (tid == 0) recs[c].e = 0;
      continue;
    }
    const uint64_t cs = k * chunk_len_of(clen[s]);  // chunk start (stream-relative)
    // A stream of blocks of one length L0 (its first block's) has its entry
    // at the first multiple of L0 at or after cs: taken when the headers
    // there and one block before both parse with length L0, without the
    // vote (a guess like the vote's; the verification decides).
    const uint32_t L0 = flen[s];
    if (L0 && L0 <= (uint32_t)kLand) {
      if (tid == 0) {
        const uint64_t g = (cs + L0 - 1) / L0 * L0;
        bool hit = g + 16 <= Ls;
        if (hit) {
          const Raw96 ra = load96(src, so, Ls, g - L0), rb = load96(src, so, Ls, g);
          uint32_t h, d1, d2;
          bool oka, okb;
          hdr96(ra, h, d1, d2);
          const uint32_t la = fast_len(h, d1, d2, &oka);
          hdr96(rb, h, d1, d2);
          const uint32_t lb = fast_len(h, d1, d2, &okb);
          hit = oka && okb && la == L0 && lb == L0;
        }
        fast_e = hit ? g : kNone;
      }
      __syncthreads();
      const uint64_t fe = fast_e;
      __syncthreads();  // (fast_e is rewritten for the next chunk)
      if (fe != kNone) {
        if (tid == 0) recs[c].e = fe;
        continue;
      }
    }
    const uint64_t ws = cs - kWarm;            // warm-up start
    const uint64_t wabs = stage<kWarm16, 256>(W, src, so + ws, so + min(cs + 32, Ls), tid);
    __syncthreads();
    const uint32_t* W32 = reinterpret_cast<const uint32_t*>(W);
    // window byte of position p0 (at most 15 + 2040: the dwords read below
    // end before byte 2072 of the 2112-byte window)
    const int b0 = (int)(so + ws - wabs) + p0;
    uint32_t cand = 0;
    {
      const int w0 = b0 >> 2;
      const uint32_t sh = (uint32_t)(b0 & 3);
      const uint32_t x0 = W32[w0], x1 = W32[w0 + 1], x2 = W32[w0 + 2], x3 = W32[w0 + 3];
      const uint32_t A0 = __builtin_amdgcn_alignbyte(x1, x0, sh);  // bytes b0 .. b0+3
      const uint32_t A1 = __builtin_amdgcn_alignbyte(x2, x1, sh);
      const uint32_t A2 = __builtin_amdgcn_alignbyte(x3, x2, sh);
#pragma unroll
      for (int u = 0; u < kVotePer; ++u) {
        const uint32_t h = __builtin_amdgcn_alignbyte(u < 4 ? A1 : A2, u < 4 ? A0 : A1, (uint32_t)(u & 3));
        cand |= hdr_ok(h) ? 1u << u : 0u;
      }
    }
    // first hop: another warm-up position, a crossing into the chunk
    // (kWarm + offset), or dead
    J4[tid] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
    uint32_t live = 0, open = 0;
    for (uint32_t m = cand; m; m &= m - 1) {
      const int u = __builtin_ctz(m);
      const int p = p0 + u;
      uint32_t h, d1, d2;
      win96(W32, b0 + u, h, d1, d2);
      // bytes to the stream end (at least 1: ws + p < cs < Ls), capped: the
      // checks compare it with lengths below 2^18
      const int64_t left = (int64_t)min(Ls - (ws + (uint64_t)p), (uint64_t)0x7FFFFFFF);
      uint32_t bl;
      const int err = step_len(h, d1, d2, left, &bl);
      // bl == left ends the stream: no entry in this chunk
      if (err == UTL_OK && (int64_t)bl < left) {
        const uint32_t q = (uint32_t)p + bl;
        if (q < (uint32_t)(kWarm + kLand)) {
          J[p] = (uint16_t)q;
          live |= 1u << u;
          if (q < (uint32_t)kWarm) open |= 1u << u;
        }
      }
    }
    __syncthreads();
    // pointer jumping over the open positions (next hop still inside the
    // warm-up): each hop doubles; (index, value) updates are single 16-bit
    // words, so a stale read is still a position on the same path. Only
    // position p's thread writes J[p].
    for (;;) {
      for (uint32_t m = open; m; m &= m - 1) {
        const int u = __builtin_ctz(m);
        const uint32_t v = J[p0 + u];  // < kWarm
        const uint32_t t = J[v];
        J[p0 + u] = (uint16_t)t;
        if (t >= (uint32_t)kWarm) open &= ~(1u << u);  // crossed, or dead (kDead)
      }
      if (!__syncthreads_or(open != 0u)) break;
    }
    for (uint32_t m = live; m; m &= m - 1) {
      const uint32_t v = J[p0 + __builtin_ctz(m)];
      if (v != kDead) atomicAdd(&hist[v - kWarm], 1u);
    }
    __syncthreads();
    // most votes; ties to the smallest offset
    uint32_t b = 0;
    for (uint32_t m = live; m; m &= m - 1) {
      const uint32_t v = J[p0 + __builtin_ctz(m)];
      if (v != kDead) b = max(b, (hist[v - kWarm] << 16) | (0xFFFFu - (v - kWarm)));
    }
    b = wave_reduce_max(b);
    if (lane_id() == 0) best[tid >> 6] = b;
    __syncthreads();
    if (tid == 0) {
      const uint32_t m = max(max(best[0], best[1]), max(best[2], best[3]));
      recs[c].e = m ? cs + (uint64_t)(0xFFFF - (m & 0xFFFF)) : kNone;
    }
    for (uint32_t m = live; m; m &= m - 1) {
      const uint32_t v = J[p0 + __builtin_ctz(m)];
      if (v != kDead) hist[v - kWarm] = 0;
    }
    __syncthreads();
  }
}

// Walk of every chunk from its voted entry.
__global__ __launch_bounds__(64) void k_walk_chunks(const uint8_t* __restrict__ src,
                                                    const uint64_t* __restrict__ stream_off, uint64_t ns,
                                                    const uint64_t* __restrict__ chunk_base, uint64_t max_chunks,
                                                    const uint32_t* __restrict__ owner_of,
                                                    const uint32_t* __restrict__ clen, WalkRec* __restrict__ recs,
                                                    uint32_t* __restrict__ pjlist, uint32_t* __restrict__ cnt) {
  // one lane per chunk: the chunks' chains of dependent header loads run
  // side by side instead of one wavefront stepping through one chunk
  const uint64_t total = min(chunk_base[ns], max_chunks);
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += step) {
    const uint64_t s = owner_of[c];
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    if (!is_long(Ls, chunk_base, s, max_chunks)) continue;
    const uint32_t cl = clen[s];
    if (cl & kClenPJ) {  // walked by k_walk_chunks_pj, which takes the chunks listed here
      pjlist[atomicAdd(&cnt[3], 1u)] = (uint32_t)c;
      continue;
    }
    const uint64_t k = c - chunk_base[s];
    const uint64_t e = recs[c].e;
    WalkRes r{kNone, 0u, kStLand};
    if (e != kNone) r = walk_lane<kCount>(src, so, Ls, e, min((k + 1) * chunk_len_of(cl), Ls), nullptr, nullptr, 0);
    recs[c].x = r.x;
    recs[c].n = r.n;
    recs[c].st = r.st;
  }
}

// Verification and block numbering of the long streams (one workgroup of
// NT threads per stream, NT chunks per round): streams of at most kSmallVer
// chunks by 64-thread workgroups, the others by 1024-thread ones, each
// class from its list (written by k_walk_plan). term[s] is the index of the
// stream's last live chunk.
// Minimum over the workgroup (NT / 64 wavefronts); red holds 32 words.
template <int NT>
__device__ __forceinline__ uint32_t wg_min(uint32_t x, uint32_t* red) {
  x = wave_reduce_min(x);
  if (NT == 64) return x;
  if (lane_id() == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t m = red[0];
  for (int w = 1; w < NT / 64; ++w) m = min(m, red[w]);
  __syncthreads();
  return m;
}

template <int NT>
__global__ __launch_bounds__(NT) void k_walk_verify(const uint8_t* __restrict__ src,
                                                    const uint64_t* __restrict__ stream_off, uint64_t ns,
                                                    const uint64_t* __restrict__ chunk_base,
                                                    const uint32_t* __restrict__ clen,
                                                    const uint32_t* __restrict__ list, const uint32_t* __restrict__ cnt,
                                                    WalkRec* __restrict__ recs, uint32_t* __restrict__ term,
                                                    uint32_t* __restrict__ nblocks, int32_t* __restrict__ status) {
  constexpr int kVer = NT;
  constexpr bool kSmall = NT == kSmallVer;
  __shared__ uint4 W[kWin16];
  __shared__ uint64_t xs[kVer];
  __shared__ uint32_t sts[kVer];
  __shared__ uint32_t red[32];
  __shared__ uint64_t fix_x;
  __shared__ uint32_t fix_n, fix_st;
  const int tid = threadIdx.x, wid = tid >> 6;
  const uint32_t nlist = cnt[kSmall ? 0 : 1];
  for (uint32_t i = blockIdx.x; i < nlist; i += gridDim.x) {
    const uint64_t s = kSmall ? list[i] : list[ns - 1 - i];
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    const uint64_t c0 = chunk_base[s];
    const uint64_t nch = chunk_base[s + 1] - c0;
    uint64_t E = 0;  // true entry of chunk k
    uint64_t k = 0;
    uint32_t total = 0, last = (uint32_t)(nch - 1), st = kStStop | UTL_ERR_INVALID_BUFFER;
    while (k < nch) {
      const uint64_t j = k + (uint64_t)tid;
      const bool in = j < nch;
      WalkRec r{kNone, kNone, 0u, kStLand, 0u, 0u};
      if (in) r = recs[c0 + j];
      xs[tid] = r.x;
      sts[tid] = r.st;
      __syncthreads();
      const uint64_t prev = tid == 0 ? E : xs[tid - 1];
      const uint32_t fb = wg_min<NT>(in && r.e != prev ? (uint32_t)tid : (uint32_t)kVer, red);
      const uint32_t ft = wg_min<NT>(in && r.st != kStLand ? (uint32_t)tid : (uint32_t)kVer, red);
      const uint32_t ngood = fb <= ft ? fb : ft + 1;  // chunks k .. k + ngood - 1 are verified
      const uint32_t cnt = (uint32_t)tid < ngood && in ? r.n : 0u;
      const uint32_t incl = wave_incl_scan(cnt);
      if (lane_id() == 63) red[wid] = incl;
      __syncthreads();
      uint32_t before = 0, sum = 0;
      for (int w = 0; w < kVer / 64; ++w) {
        before += w < wid ? red[w] : 0u;
        sum += red[w];
      }
      if ((uint32_t)tid < ngood && in) recs[c0 + j].base = total + before + incl - cnt;
      total += sum;
      if (ft < (uint32_t)kVer && ft < fb) {  // verified up to a stop or the stream end
        last = (uint32_t)(k + ft);
        st = sts[ft];
        __syncthreads();
        break;
      }
      if (fb == (uint32_t)kVer) {  // all verified, none final
        const bool more = nch - k > (uint64_t)kVer;
        E = xs[kVer - 1];
        k += kVer;
        __syncthreads();
        if (!more) break;  // cannot happen: a stream's last chunk always ends the walk
        continue;
      }
      // chunk k + fb has a wrong entry: walk it again from the true one
      E = fb == 0 ? E : xs[fb - 1];
      k += fb;
      if (wid == 0) {
        const WalkRes w = walk_chunk(W, src, so, Ls, chunk_len_of(clen[s]), k, E, nullptr, nullptr, 0);
        if (lane_id() == 0) {
          WalkRec f{E, w.x, w.n, w.st, total, 0u};
          recs[c0 + k] = f;
          fix_x = w.x;
          fix_n = w.n;
          fix_st = w.st;
        }
      }
      __syncthreads();
      total += fix_n;
      const uint32_t fst = fix_st;
      E = fix_x;
      __syncthreads();
      if (fst != kStLand) {
        last = (uint32_t)k;
        st = fst;
        break;
      }
      ++k;
    }
    if (tid == 0) {
      term[s] = last;
      nblocks[s] = total;
      status[s] = status_of(st);
    }
  }
}

// Meta words and block starts of every live chunk (one lane per chunk).
__global__ __launch_bounds__(64) void k_walk_emit(const uint8_t* __restrict__ src,
                                                  const uint64_t* __restrict__ stream_off, uint64_t ns,
                                                  const uint64_t* __restrict__ chunk_base, uint64_t max_chunks,
                                                  const uint32_t* __restrict__ owner_of,
                                                  const uint32_t* __restrict__ clen, const WalkRec* __restrict__ recs,
                                                  const uint32_t* __restrict__ term,
                                                  const uint64_t* __restrict__ blk_off, uint32_t* __restrict__ meta,
                                                  uint64_t* __restrict__ start) {
  const uint64_t total = min(chunk_base[ns], max_chunks);
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += step) {
    const uint64_t s = owner_of[c];
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    if (!is_long(Ls, chunk_base, s, max_chunks)) continue;
    const uint32_t cl = clen[s];
    if (cl & kClenPJ) continue;  // k_walk_chunks_pj
    const uint64_t k = c - chunk_base[s];
    if (k > term[s]) continue;
    const WalkRec r = recs[c];
    if (r.n == 0) continue;
    walk_lane_out(src, so, Ls, r.e, min((k + 1) * chunk_len_of(cl), Ls), meta, start, blk_off[s] + r.base);
  }
}

// Serial walk of the short streams (one lane per stream); with a plan
// (chunk_base set), not the long streams nor the dense short ones.
__global__ __launch_bounds__(256) void k_walk_serial(const uint8_t* __restrict__ src,
                                                     const uint64_t* __restrict__ stream_off, uint64_t ns,
                                                     const uint64_t* __restrict__ chunk_base, uint64_t max_chunks,
                                                     const uint32_t* __restrict__ clen,
                                                     const uint64_t* __restrict__ blk_off, uint32_t* __restrict__ meta,
                                                     uint64_t* __restrict__ start, uint32_t* __restrict__ nblocks,
                                                     int32_t* __restrict__ status) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += step) {
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    if (chunk_base && (is_long(Ls, chunk_base, s, max_chunks) || (Ls < kLong && clen[s] == kDense))) continue;
    const WalkRes r = walk_lane_out(src, so, Ls, 0, Ls, meta, start, meta ? blk_off[s] : 0);
    nblocks[s] = r.n;
    status[s] = status_of(r.st);
  }
}

// Window walk by pointer jumping (a workgroup of kPJThreads threads). Every
// position p of a window of kPJ stream positions gets the position of the
// block after a block starting at p (J0; kJOut when that lies at or past
// the window's limit, kJEnd when the block ends the stream, kJDead when no
// block can start at p: the reference's checks fail or the block is longer
// than the bytes left). From the window's first position, a block start,
// round k marks the chain positions at distances [2^k, 2^(k+1)) through the
// table of 2^k jumps and builds the table of 2^(k+1) jumps (tables hold only
// chain positions or terminals, and every jump moves forward, so a mark set
// early in a round is still a chain position and the rounds end). When the
// first position jumps to a terminal, every chain position is marked; the
// marked positions other than a kJDead one are the window's blocks, in
// order, and the last marked position tells how the walk goes on (the next
// window, the walk's end, the stream end or a stop).
constexpr int kPJ = 4096;                 // positions of a window
constexpr int kPJ16 = (kPJ + 32) / 16;    // its staged bytes (12 header bytes past the last position)
constexpr int kPJThreads = 256;
constexpr int kPJPer = kPJ / kPJThreads;  // consecutive positions per thread
constexpr uint16_t kJOut = kPJ, kJEnd = 0xFFFE, kJDead = 0xFFFF;
static_assert(kPJ16 * 16 >= kPJ + 16 + 16, "a window holds 12 header bytes past its last position");

struct PJLds {
  uint4 W[kPJ16];
  uint16_t J0[kPJ], JA[kPJ], JB[kPJ];
  uint8_t mark[kPJ];
  uint32_t red[kPJThreads / 64];
  uint32_t last_p;
};

// BlockLength at window byte i (stream bytes left from it: left >= 1) as
// the lane walk checks it; err is the stop status when no block can start
// there (UTL_OK otherwise).
__device__ __forceinline__ uint32_t pj_len(const uint32_t* W32, int i, uint64_t left, int& err, uint32_t& h) {
  uint32_t d1, d2;
  win96(W32, i, h, d1, d2);
  bool ok;
  uint32_t bl = fast_len(h, d1, d2, &ok);
  err = UTL_OK;
  if (!ok || left < 16u) err = step_len(h, d1, d2, (int64_t)min(left, (uint64_t)0x7FFFFFFF), &bl);
  if (err == UTL_OK && bl > left) err = UTL_ERR_INVALID_BUFFER;
  return bl;
}

// walk_lane's walk (stream at absolute so, length Ls) from the block start
// pos while pos < end, by windows; with meta set, block i's meta word goes
// to meta[out + i] and, with start set, its absolute start to start[out +
// i]. Every thread of the workgroup calls it with the same arguments.
__device__ WalkRes pj_walk(PJLds& L, const uint8_t* __restrict__ src, uint64_t so, uint64_t Ls, uint64_t pos,
                           uint64_t end, uint32_t* __restrict__ meta, uint64_t* __restrict__ start, uint64_t out) {
  const int tid = threadIdx.x, wid = tid >> 6;
  const int p0 = tid * kPJPer;
  uint32_t n = 0, st = kStLand;
  while (pos < end) {
    const int lim = (int)min(end - pos, (uint64_t)kPJ);  // positions of this window
    const uint64_t wabs = stage<kPJ16, kPJThreads>(L.W, src, so + pos, so + min(pos + kPJ + 16, Ls), tid);
    __syncthreads();
    const uint32_t* W32 = reinterpret_cast<const uint32_t*>(L.W);
    const int a0 = (int)(so + pos - wabs);
#pragma unroll 4
    for (int u = 0; u < kPJPer; ++u) {
      const int p = p0 + u;
      uint16_t j = kJDead;
      if (p < lim) {  // (p < end - pos <= Ls - pos)
        const uint64_t left = Ls - pos - (uint64_t)p;
        int err;
        uint32_t h;
        const uint32_t bl = pj_len(W32, a0 + p, left, err, h);
        if (err == UTL_OK) j = bl == left ? kJEnd : (uint32_t)p + bl >= (uint32_t)lim ? kJOut : (uint16_t)(p + bl);
      }
      L.J0[p] = j;
      L.JA[p] = j;
      L.mark[p] = p == 0 ? 1 : 0;
    }
    __syncthreads();
    uint16_t* cur = L.JA;
    uint16_t* nxt = L.JB;
    for (;;) {
#pragma unroll 4
      for (int u = 0; u < kPJPer; ++u) {
        const int p = p0 + u;
        const uint16_t t = cur[p];
        if (L.mark[p] && t < kPJ) L.mark[t] = 1;
        nxt[p] = t < kPJ ? cur[t] : t;
      }
      __syncthreads();
      const bool done = nxt[0] >= kPJ;
      uint16_t* tmp = cur;
      cur = nxt;
      nxt = tmp;
      if (done) break;
    }
    // the window's blocks: marked positions other than a kJDead one; the
    // last marked position has a terminal J0
    uint32_t c = 0;
#pragma unroll 4
    for (int u = 0; u < kPJPer; ++u) {
      const int p = p0 + u;
      if (L.mark[p]) {
        const uint16_t j = L.J0[p];
        c += j != kJDead ? 1u : 0u;
        if (j >= kPJ) L.last_p = (uint32_t)p;
      }
    }
    const uint32_t incl = wave_incl_scan(c);
    if (lane_id() == 63) L.red[wid] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kPJThreads / 64; ++w) {
      before += w < wid ? L.red[w] : 0u;
      total += L.red[w];
    }
    if (meta) {
      uint64_t k = out + n + before + incl - c;  // this thread's first block
      for (int u = 0; u < kPJPer; ++u) {
        const int p = p0 + u;
        if (L.mark[p] && L.J0[p] != kJDead) {
          int err;
          uint32_t h;
          const uint32_t bl = pj_len(W32, a0 + p, Ls - pos - (uint64_t)p, err, h);
          meta[k] = (bl & 0xFFFFFFu) | ((h & 0xFFu) << 24);
          if (start) start[k] = so + pos + (uint64_t)p;
          ++k;
        }
      }
    }
    n += total;
    const int lp = (int)L.last_p;
    const uint16_t lj = L.J0[lp];
    int err;
    uint32_t h;
    const uint32_t bl = pj_len(W32, a0 + lp, Ls - pos - (uint64_t)lp, err, h);
    __syncthreads();  // W, the tables, the marks and red are rewritten by the next window
    if (lj == kJDead) {  // (the walk stops at the position that is not a block, as walk_lane's)
      pos += (uint64_t)lp;
      st = kStStop | (uint32_t)err;
      break;
    }
    if (lj == kJEnd) {
      pos = Ls;
      st = kStEnd;
      break;
    }
    pos += (uint64_t)lp + bl;  // kJOut: the block crossing the window's limit
  }
  return {pos, n, st};
}

// Walk of the dense short streams of dlist, one workgroup per stream.
__global__ __launch_bounds__(kPJThreads) void k_walk_dense(const uint8_t* __restrict__ src,
                                                           const uint64_t* __restrict__ stream_off,
                                                           const uint32_t* __restrict__ dlist,
                                                           const uint32_t* __restrict__ cnt,
                                                           const uint64_t* __restrict__ blk_off,
                                                           uint32_t* __restrict__ meta, uint64_t* __restrict__ start,
                                                           uint32_t* __restrict__ nblocks,
                                                           int32_t* __restrict__ status) {
  __shared__ PJLds L;
  const uint32_t nd = cnt[2];
  for (uint32_t i = blockIdx.x; i < nd; i += gridDim.x) {
    const uint64_t s = dlist[i];
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    const WalkRes r = pj_walk(L, src, so, Ls, 0, Ls, meta, meta ? start : nullptr, meta ? blk_off[s] : 0);
    if (threadIdx.x == 0) {
      nblocks[s] = r.n;
      status[s] = status_of(r.st);
    }
  }
}

// The chunks of the streams flagged kClenPJ (listed by k_walk_chunks in
// pjlist, cnt[3] of them), one workgroup per chunk: with meta unset,
// k_walk_chunks' walk from the voted entry (block count, exit, stop status
// into the chunk's record); with meta set, k_walk_emit's (the live chunks'
// meta words and starts at their verified bases).
__global__ __launch_bounds__(kPJThreads) void k_walk_chunks_pj(
    const uint8_t* __restrict__ src, const uint64_t* __restrict__ stream_off, uint64_t ns,
    const uint64_t* __restrict__ chunk_base, uint64_t max_chunks, const uint32_t* __restrict__ owner_of,
    const uint32_t* __restrict__ clen, const uint32_t* __restrict__ pjlist, const uint32_t* __restrict__ cnt,
    WalkRec* __restrict__ recs,
    const uint32_t* __restrict__ term, const uint64_t* __restrict__ blk_off, uint32_t* __restrict__ meta,
    uint64_t* __restrict__ start) {
  __shared__ PJLds L;
  const uint32_t np = cnt[3];
  for (uint32_t i = blockIdx.x; i < np; i += gridDim.x) {
    const uint64_t c = pjlist[i];
    const uint64_t s = owner_of[c];
    uint64_t so;
    const uint64_t Ls = stream_len(stream_off, s, &so);
    if (!is_long(Ls, chunk_base, s, max_chunks)) continue;
    const uint32_t cl = clen[s];
    if (!(cl & kClenPJ)) continue;
    const uint64_t k = c - chunk_base[s];
    const uint64_t end = min((k + 1) * chunk_len_of(cl), Ls);
    if (!meta) {
      const uint64_t e = recs[c].e;
      WalkRes r{kNone, 0u, kStLand};
      if (e != kNone) r = pj_walk(L, src, so, Ls, e, end, nullptr, nullptr, 0);
      if (threadIdx.x == 0) {
        recs[c].x = r.x;
        recs[c].n = r.n;
        recs[c].st = r.st;
      }
    } else {
      if (k > term[s]) continue;
      const WalkRec r = recs[c];
      if (r.n == 0) continue;
      pj_walk(L, src, so, Ls, r.e, end, meta, start, blk_off[s] + r.base);
    }
  }
}

}  // namespace
}  // namespace utl

using namespace utl;

// Chunk length of the long-stream walk for a walk over total_bytes. Every
// chunk costs one entry vote (about 6 ns of the whole device at full
// occupancy) and the walk of a chunk takes about 2.6 ns per byte (one
// dependent header load per block at 234-byte blocks), so the sum for B
// bytes is least near sqrt(2.35 B). The length stays within
// [kChunk, 4 kChunk]: longer chunks of small blocks make long chains.
static uint64_t walk_chunk_len(uint64_t total_bytes) {
  const uint64_t c = (uint64_t)std::sqrt(2.35 * (double)total_bytes);
  return std::min(std::max(c, (uint64_t)kChunk), (uint64_t)4 * kChunk) & ~15ull;
}

// Scratch layout: chunk_base (nstreams + 1 entries), tile sums, chunk
// records (at most one per kChunk bytes plus one per long stream: a chunk
// is never shorter than kChunk), term, the long-stream lists, the chunk
// lengths, the first block lengths and the dense-stream list (one word per
// stream each), the chunks' streams and the list of chunks walked by
// workgroups (one word per chunk each), four list counters.
static uint64_t walk_words(uint64_t n, uint64_t bytes) { return (n * bytes + 255) / 256 * 256; }

size_t utl_walk_tmp_bytes(uint64_t nstreams, uint64_t total_bytes) {
  const uint64_t max_chunks = total_bytes / kChunk + total_bytes / kLong + 1;
  const uint64_t ntiles = (nstreams + 1 + kPlan - 1) / kPlan;
  return walk_words(nstreams + 1, 8) + walk_words(ntiles, 8) + max_chunks * sizeof(WalkRec) +
         5 * walk_words(nstreams, 4) + 2 * walk_words(max_chunks, 4) + 256;
}

hipError_t utl_launch_index_walk(const uint8_t* src, const uint64_t* stream_off, uint64_t nstreams,
                                 uint64_t total_bytes, const uint64_t* blk_off, uint32_t* meta, uint64_t* start,
                                 uint32_t* nblocks, int32_t* status, void* tmp, hipStream_t stream,
                                 hipStream_t side, hipEvent_t ev_fork, hipEvent_t ev_join) {
  if (nstreams == 0) return hipSuccess;
  if (nstreams > 0xFFFFFFFFull) return hipErrorInvalidValue;  // stream numbers are 32-bit words
  const uint64_t max_chunks = total_bytes / kChunk + total_bytes / kLong + 1;
  const uint64_t chunk = walk_chunk_len(total_bytes);
  const uint64_t ntiles = (nstreams + 1 + kPlan - 1) / kPlan;
  uint8_t* t = static_cast<uint8_t*>(tmp);
  uint64_t* chunk_base = reinterpret_cast<uint64_t*>(t);
  t += walk_words(nstreams + 1, 8);
  uint64_t* tile_sum = reinterpret_cast<uint64_t*>(t);
  t += walk_words(ntiles, 8);
  WalkRec* recs = reinterpret_cast<WalkRec*>(t);
  t += max_chunks * sizeof(WalkRec);
  uint32_t* term = reinterpret_cast<uint32_t*>(t);
  t += walk_words(nstreams, 4);
  uint32_t* list = reinterpret_cast<uint32_t*>(t);
  t += walk_words(nstreams, 4);
  uint32_t* clen = reinterpret_cast<uint32_t*>(t);
  t += walk_words(nstreams, 4);
  uint32_t* flen = reinterpret_cast<uint32_t*>(t);
  t += walk_words(nstreams, 4);
  uint32_t* dlist = reinterpret_cast<uint32_t*>(t);
  t += walk_words(nstreams, 4);
  uint32_t* pjlist = reinterpret_cast<uint32_t*>(t);
  t += walk_words(max_chunks, 4);
  uint32_t* owner_of = reinterpret_cast<uint32_t*>(t);
  t += walk_words(max_chunks, 4);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(t);
  // the plan (chunk lengths, dense short streams) pays from a few KiB on
  const bool planned = total_bytes >= 4096;
  if (planned) {
    k_walk_zero<<<dim3(1), dim3(4), 0, stream>>>(cnt);
    k_walk_plan<<<dim3((unsigned)ntiles), dim3(kPlan), 0, stream>>>(src, stream_off, nstreams, chunk, chunk_base,
                                                                     tile_sum, clen, flen, list, dlist, cnt);
    if (ntiles > 1) {
      k_walk_scan<<<dim3(1), dim3(1024), 0, stream>>>(tile_sum, ntiles);
      k_walk_apply<<<dim3((unsigned)(ntiles - 1)), dim3(kPlan), 0, stream>>>(chunk_base, nstreams, tile_sum);
    }
  }
  // The short streams (lane walks and dense streams) need only the plan:
  // with many streams of a mean length that suggests long ones as well,
  // they run on the side stream beside the long streams' vote, chunk walks,
  // verification and emit, and the context's stream waits for them at the
  // end (the fork and join cost about 20 us, more than a walk of a few
  // streams gains).
  const bool fork = planned && nstreams >= 1024 && total_bytes / nstreams >= kLong;
  const hipStream_t sh = fork ? side : stream;  // the short streams' stream
  if (fork) {
    const hipError_t e = hipEventRecord(ev_fork, stream);
    if (e != hipSuccess) return e;
    const hipError_t e2 = hipStreamWaitEvent(side, ev_fork, 0);
    if (e2 != hipSuccess) return e2;
  }
  if (planned) {
    // dense short streams: at most one per kDenseChain * 4 + 1 bytes (a block is at least 4 bytes)
    const uint64_t ndense_max = std::min(nstreams, total_bytes / (kDenseChain * 4 + 1) + 1);
    k_walk_dense<<<dim3((unsigned)std::min(ndense_max, (uint64_t)1024)), dim3(kPJThreads), 0, sh>>>(
        src, stream_off, dlist, cnt, blk_off, meta, start, nblocks, status);
    const unsigned lsgrid = (unsigned)std::min((nstreams + 255) / 256, (uint64_t)4096);  // one lane per stream
    k_walk_serial<<<dim3(lsgrid), dim3(256), 0, sh>>>(src, stream_off, nstreams, chunk_base, max_chunks, clen,
                                                      blk_off, meta, start, nblocks, status);
  }
  if (fork) {
    const hipError_t e = hipEventRecord(ev_join, side);
    if (e != hipSuccess) return e;
  }
  if (planned && total_bytes >= kLong) {
    k_walk_owner<<<dim3((unsigned)ntiles), dim3(kPlan), 0, stream>>>(chunk_base, nstreams, max_chunks, owner_of);
    // grids from the bound on chunks (their count is on the device): every
    // chunk has its own lane; workgroups past the count exit at once
    const unsigned vgrid = (unsigned)std::min(max_chunks, (uint64_t)4096);
    k_walk_vote<<<dim3(vgrid), dim3(256), 0, stream>>>(src, stream_off, nstreams, chunk_base, max_chunks, owner_of,
                                                       clen, flen, recs);
    const unsigned lgrid = (unsigned)std::min((max_chunks + 63) / 64, (uint64_t)16384);  // one lane per chunk
    k_walk_chunks<<<dim3(lgrid), dim3(64), 0, stream>>>(src, stream_off, nstreams, chunk_base, max_chunks, owner_of,
                                                        clen, recs, pjlist, cnt);
    // the chunks of small-block streams: a workgroup each (their count is on
    // the device; 1024 workgroups are resident at once)
    const unsigned pjgrid = (unsigned)std::min(max_chunks, (uint64_t)1024);
    k_walk_chunks_pj<<<dim3(pjgrid), dim3(kPJThreads), 0, stream>>>(src, stream_off, nstreams, chunk_base,
                                                                    max_chunks, owner_of, clen, pjlist, cnt, recs,
                                                                    nullptr, nullptr, nullptr, nullptr);
    // grids: every workgroup resident at once (a stream's verification is
    // a few rounds of loads and barriers)
    const uint64_t nlong_max = total_bytes / kLong + 1;
    k_walk_verify<kSmallVer><<<dim3((unsigned)std::min(nlong_max, (uint64_t)8192)), dim3(kSmallVer), 0, stream>>>(
        src, stream_off, nstreams, chunk_base, clen, list, cnt, recs, term, nblocks, status);
    k_walk_verify<1024><<<dim3((unsigned)std::min(nlong_max, (uint64_t)512)), dim3(1024), 0, stream>>>(
        src, stream_off, nstreams, chunk_base, clen, list, cnt, recs, term, nblocks, status);
    if (meta) {
      k_walk_emit<<<dim3(lgrid), dim3(64), 0, stream>>>(src, stream_off, nstreams, chunk_base, max_chunks,
                                                        owner_of, clen, recs, term, blk_off, meta, start);
      k_walk_chunks_pj<<<dim3(pjgrid), dim3(kPJThreads), 0, stream>>>(src, stream_off, nstreams, chunk_base,
                                                                      max_chunks, owner_of, clen, pjlist, cnt, recs,
                                                                      term, blk_off, meta, start);
    }
  }
  if (fork) {
    const hipError_t e = hipStreamWaitEvent(stream, ev_join, 0);
    if (e != hipSuccess) return e;
  } else if (!planned) {
    const unsigned lsgrid = (unsigned)std::min((nstreams + 255) / 256, (uint64_t)4096);  // one lane per stream
    k_walk_serial<<<dim3(lsgrid), dim3(256), 0, stream>>>(src, stream_off, nstreams, nullptr, max_chunks, clen,
                                                          blk_off, meta, start, nblocks, status);
  }
  return hipGetLastError();
}
