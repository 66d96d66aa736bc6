// utl_index.hip -- stream index kernels.
//
// The reference walks a concatenated stream serially: block k+1 starts where
// BlockLength (block_length.go:8) of block k ends (example_test.go:171-187).
// Here the per-block meta words (byte length | value count << 24) that the
// encoders emit are turned into block byte offsets and output value offsets
// by a device exclusive scan (tile reduce, tile-prefix scan, tile apply; the
// middle pass is folded into the apply pass when there are few tiles),
// which removes the loop-carried dependency. A header-parse pass
// (k_block_lengths) recomputes BlockLength for every block whose start is
// known, on the device.
#include "utl_device.h"
#include "utl_format.h"
#include "utl_kernels.h"

namespace utl {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kTile = kScanThreads * kScanItems;

__device__ __forceinline__ void split_meta(uint32_t m, uint64_t& len, uint64_t& cnt) {
  len = m & 0xFFFFFFu;
  cnt = m >> 24;
}

// Exclusive scan over a 256-thread workgroup. lds holds 4 entries.
__device__ __forceinline__ uint64_t wg_excl_scan(uint64_t x, uint64_t* lds, uint64_t& total) {
  const int tid = threadIdx.x;
  const uint64_t incl = wave_incl_scan64(x);
  if ((tid & 63) == 63) lds[tid >> 6] = incl;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    const uint64_t v = lds[w];
    if (w < (tid >> 6)) pre += v;
    tot += v;
  }
  __syncthreads();
  total = tot;
  return incl - x + pre;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ meta, uint64_t n,
                                                               uint64_t* __restrict__ tile_len,
                                                               uint64_t* __restrict__ tile_cnt) {
  __shared__ uint64_t lds[2][4];
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t sl = 0, sc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) {
      uint64_t l, c;
      split_meta(meta[base + i], l, c);
      sl += l;
      sc += c;
    }
  }
  uint64_t tl, tc;
  wg_excl_scan(sl, lds[0], tl);
  wg_excl_scan(sc, lds[1], tc);
  if (threadIdx.x == 0) {
    tile_len[blockIdx.x] = tl;
    tile_cnt[blockIdx.x] = tc;
  }
}

// One workgroup turns tile totals into exclusive tile prefixes, in place.
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(uint64_t* __restrict__ tile_len,
                                                              uint64_t* __restrict__ tile_cnt, uint64_t ntiles) {
  __shared__ uint64_t lds[2][4];
  uint64_t carry_l = 0, carry_c = 0;
  for (uint64_t base = 0; base < ntiles; base += kScanThreads) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t l = i < ntiles ? tile_len[i] : 0, c = i < ntiles ? tile_cnt[i] : 0;
    uint64_t tl, tc;
    const uint64_t el = wg_excl_scan(l, lds[0], tl);
    const uint64_t ec = wg_excl_scan(c, lds[1], tc);
    if (i < ntiles) {
      tile_len[i] = carry_l + el;
      tile_cnt[i] = carry_c + ec;
    }
    carry_l += tl;
    carry_c += tc;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* __restrict__ meta, uint64_t n,
                                                              const uint64_t* __restrict__ tile_len,
                                                              const uint64_t* __restrict__ tile_cnt,
                                                              uint64_t* __restrict__ boff, uint64_t* __restrict__ ooff) {
  __shared__ uint64_t lds[2][4];
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t l[kScanItems], c[kScanItems], sl = 0, sc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    l[i] = c[i] = 0;
    if (base + i < n) split_meta(meta[base + i], l[i], c[i]);
    sl += l[i];
    sc += c[i];
  }
  uint64_t tl, tc;
  uint64_t pl = wg_excl_scan(sl, lds[0], tl) + tile_len[blockIdx.x];
  uint64_t pc = wg_excl_scan(sc, lds[1], tc) + tile_cnt[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) {
      boff[base + i] = pl;
      ooff[base + i] = pc;
    }
    pl += l[i];
    pc += c[i];
    if (base + i + 1 == n) {
      boff[n] = pl;
      ooff[n] = pc;
    }
  }
}

// Coalesced form (16-byte aligned meta, block_off and out_off): tiles of
// 1024 blocks, thread t owning blocks 4t .. 4t+3 of its tile, read with one
// 16-byte load and written with 16-byte stores. With at most kInlineTiles
// tiles the apply pass sums the earlier tiles' totals itself, which saves
// the tile-prefix pass.
constexpr int kTile4 = kScanThreads * 4;
constexpr uint64_t kInlineTiles = 1024;

__device__ __forceinline__ void load_meta4(const uint32_t* __restrict__ meta, uint64_t n, uint64_t i,
                                           uint64_t* l, uint64_t* c) {
  uint32_t m[4] = {0u, 0u, 0u, 0u};
  if (i + 4 <= n) {
    const uint4 q = *reinterpret_cast<const uint4*>(meta + i);
    m[0] = q.x;
    m[1] = q.y;
    m[2] = q.z;
    m[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + k < n) m[k] = meta[i + k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) split_meta(m[k], l[k], c[k]);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce4(const uint32_t* __restrict__ meta, uint64_t n,
                                                                uint64_t* __restrict__ tile_len,
                                                                uint64_t* __restrict__ tile_cnt) {
  __shared__ uint64_t lds[2][4];
  uint64_t l[4], c[4];
  load_meta4(meta, n, (uint64_t)blockIdx.x * kTile4 + 4 * threadIdx.x, l, c);
  uint64_t tl, tc;
  wg_excl_scan(l[0] + l[1] + l[2] + l[3], lds[0], tl);
  wg_excl_scan(c[0] + c[1] + c[2] + c[3], lds[1], tc);
  if (threadIdx.x == 0) {
    tile_len[blockIdx.x] = tl;
    tile_cnt[blockIdx.x] = tc;
  }
}

// Groups of kGroupTiles tiles: k_scan_groups turns the tile totals of each
// group into exclusive prefixes within the group, in place, and writes the
// group's total; the apply pass then adds the totals of the groups before
// its own (at most kInlineTiles of them, summed by the workgroup itself).
constexpr uint64_t kGroupTiles = kTile4;

__global__ __launch_bounds__(kScanThreads) void k_scan_groups(uint64_t* __restrict__ tile_len,
                                                               uint64_t* __restrict__ tile_cnt, uint64_t ntiles,
                                                               uint64_t* __restrict__ group_len,
                                                               uint64_t* __restrict__ group_cnt) {
  __shared__ uint64_t lds[2][4];
  const uint64_t i0 = (uint64_t)blockIdx.x * kGroupTiles + 4 * threadIdx.x;
  uint64_t l[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    l[k] = i0 + k < ntiles ? tile_len[i0 + k] : 0;
    c[k] = i0 + k < ntiles ? tile_cnt[i0 + k] : 0;
  }
  uint64_t tl, tc;
  uint64_t pl = wg_excl_scan(l[0] + l[1] + l[2] + l[3], lds[0], tl);
  uint64_t pc = wg_excl_scan(c[0] + c[1] + c[2] + c[3], lds[1], tc);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (i0 + k < ntiles) {
      tile_len[i0 + k] = pl;
      tile_cnt[i0 + k] = pc;
    }
    pl += l[k];
    pc += c[k];
  }
  if (threadIdx.x == 0) {
    group_len[blockIdx.x] = tl;
    group_cnt[blockIdx.x] = tc;
  }
}

// MODE 0: tile_len/tile_cnt hold tile totals and each workgroup sums those
// before it. MODE 1: they hold prefixes within groups of kGroupTiles tiles
// (k_scan_groups) and each workgroup sums the group totals before its own.
// MODE 2: as 1, with the group totals already turned into exclusive group
// prefixes (k_scan_tiles over the groups).
template <int MODE>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply4(const uint32_t* __restrict__ meta, uint64_t n,
                                                               const uint64_t* __restrict__ tile_len,
                                                               const uint64_t* __restrict__ tile_cnt,
                                                               const uint64_t* __restrict__ group_len,
                                                               const uint64_t* __restrict__ group_cnt,
                                                               uint64_t* __restrict__ boff,
                                                               uint64_t* __restrict__ ooff) {
  __shared__ uint64_t lds[4][4];
  const uint64_t i0 = (uint64_t)blockIdx.x * kTile4 + 4 * threadIdx.x;
  uint64_t l[4], c[4];
  load_meta4(meta, n, i0, l, c);
  uint64_t tile_pl, tile_pc;
  if (MODE == 0) {
    uint64_t sl = 0, sc = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) {
      sl += tile_len[t];
      sc += tile_cnt[t];
    }
    wg_excl_scan(sl, lds[2], tile_pl);
    wg_excl_scan(sc, lds[3], tile_pc);
  } else if (MODE == 1) {
    const uint32_t grp = (uint32_t)(blockIdx.x / kGroupTiles);
    uint64_t sl = 0, sc = 0;
    for (uint32_t t = threadIdx.x; t < grp; t += kScanThreads) {
      sl += group_len[t];
      sc += group_cnt[t];
    }
    wg_excl_scan(sl, lds[2], tile_pl);
    wg_excl_scan(sc, lds[3], tile_pc);
    tile_pl += tile_len[blockIdx.x];
    tile_pc += tile_cnt[blockIdx.x];
  } else {
    const uint64_t grp = blockIdx.x / kGroupTiles;
    tile_pl = tile_len[blockIdx.x] + group_len[grp];
    tile_pc = tile_cnt[blockIdx.x] + group_cnt[grp];
  }
  uint64_t tl, tc;
  uint64_t pl = wg_excl_scan(l[0] + l[1] + l[2] + l[3], lds[0], tl) + tile_pl;
  uint64_t pc = wg_excl_scan(c[0] + c[1] + c[2] + c[3], lds[1], tc) + tile_pc;
  uint64_t b[5], o[5];
  b[0] = pl;
  o[0] = pc;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    b[k + 1] = b[k] + l[k];
    o[k + 1] = o[k] + c[k];
  }
  if (i0 + 4 <= n) {
    *reinterpret_cast<ulonglong2*>(boff + i0) = make_ulonglong2(b[0], b[1]);
    *reinterpret_cast<ulonglong2*>(boff + i0 + 2) = make_ulonglong2(b[2], b[3]);
    *reinterpret_cast<ulonglong2*>(ooff + i0) = make_ulonglong2(o[0], o[1]);
    *reinterpret_cast<ulonglong2*>(ooff + i0 + 2) = make_ulonglong2(o[2], o[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i0 + k < n) {
        boff[i0 + k] = b[k];
        ooff[i0 + k] = o[k];
      }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (i0 + k + 1 == n) {
      boff[n] = b[k + 1];
      ooff[n] = o[k + 1];
    }
}

// The wave index in tiles of kWaveTile blocks: thread t owns blocks 16t ..
// 16t + 15 of its tile (four 16-byte meta loads in flight) and writes the
// entries of its two groups of eight (32 contiguous bytes). Four times the
// blocks per thread of the block-index kernels, for more bytes in flight per
// wave and a quarter of the tiles; the passes and MODEs are theirs.
constexpr int kWaveItems = 16;
constexpr uint64_t kWaveTile = (uint64_t)kScanThreads * kWaveItems;

__device__ __forceinline__ void load_meta16(const uint32_t* __restrict__ meta, uint64_t n, uint64_t i,
                                            uint32_t* m) {
#pragma unroll
  for (int q = 0; q < kWaveItems / 4; ++q) {
    const uint64_t k = i + 4 * q;
    if (k + 4 <= n) {
      const uint4 v = *reinterpret_cast<const uint4*>(meta + k);
      m[4 * q] = v.x;
      m[4 * q + 1] = v.y;
      m[4 * q + 2] = v.z;
      m[4 * q + 3] = v.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) m[4 * q + e] = k + e < n ? meta[k + e] : 0u;
    }
  }
}

__global__ __launch_bounds__(kScanThreads) void k_wave_reduce(const uint32_t* __restrict__ meta, uint64_t n,
                                                               uint64_t* __restrict__ tile_len,
                                                               uint64_t* __restrict__ tile_cnt) {
  __shared__ uint64_t lds[2][4];
  uint32_t m[kWaveItems];
  load_meta16(meta, n, (uint64_t)blockIdx.x * kWaveTile + kWaveItems * threadIdx.x, m);
  uint64_t sl = 0, sc = 0;
#pragma unroll
  for (int k = 0; k < kWaveItems; ++k) {
    sl += m[k] & 0xFFFFFFu;
    sc += m[k] >> 24;
  }
  uint64_t tl, tc;
  wg_excl_scan(sl, lds[0], tl);
  wg_excl_scan(sc, lds[1], tc);
  if (threadIdx.x == 0) {
    tile_len[blockIdx.x] = tl;
    tile_cnt[blockIdx.x] = tc;
  }
}

template <int MODE>
__global__ __launch_bounds__(kScanThreads) void k_wave_apply(const uint32_t* __restrict__ meta, uint64_t n,
                                                              const uint64_t* __restrict__ tile_len,
                                                              const uint64_t* __restrict__ tile_cnt,
                                                              const uint64_t* __restrict__ group_len,
                                                              const uint64_t* __restrict__ group_cnt,
                                                              uint64_t* __restrict__ wave) {
  __shared__ uint64_t lds[4][4];
  const uint64_t i0 = (uint64_t)blockIdx.x * kWaveTile + kWaveItems * threadIdx.x;
  uint32_t m[kWaveItems];
  load_meta16(meta, n, i0, m);
  uint64_t tile_pl, tile_pc;
  if (MODE == 0) {
    uint64_t sl = 0, sc = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) {
      sl += tile_len[t];
      sc += tile_cnt[t];
    }
    wg_excl_scan(sl, lds[2], tile_pl);
    wg_excl_scan(sc, lds[3], tile_pc);
  } else if (MODE == 1) {
    const uint32_t grp = (uint32_t)(blockIdx.x / kGroupTiles);
    uint64_t sl = 0, sc = 0;
    for (uint32_t t = threadIdx.x; t < grp; t += kScanThreads) {
      sl += group_len[t];
      sc += group_cnt[t];
    }
    wg_excl_scan(sl, lds[2], tile_pl);
    wg_excl_scan(sc, lds[3], tile_pc);
    tile_pl += tile_len[blockIdx.x];
    tile_pc += tile_cnt[blockIdx.x];
  } else {
    const uint64_t grp = blockIdx.x / kGroupTiles;
    tile_pl = tile_len[blockIdx.x] + group_len[grp];
    tile_pc = tile_cnt[blockIdx.x] + group_cnt[grp];
  }
  uint64_t l8 = 0, c8 = 0, sl = 0, sc = 0;
#pragma unroll
  for (int k = 0; k < kWaveItems; ++k) {
    if (k == 8) {
      l8 = sl;
      c8 = sc;
    }
    sl += m[k] & 0xFFFFFFu;
    sc += m[k] >> 24;
  }
  uint64_t tl, tc;
  const uint64_t pl = wg_excl_scan(sl, lds[0], tl) + tile_pl;
  const uint64_t pc = wg_excl_scan(sc, lds[1], tc) + tile_pc;
  uint64_t* w = wave + 2 * (i0 >> 3);
  if (i0 + 8 < n) {
    *reinterpret_cast<ulonglong2*>(w) = make_ulonglong2(pl, pc);
    *reinterpret_cast<ulonglong2*>(w + 2) = make_ulonglong2(pl + l8, pc + c8);
  } else if (i0 < n) {
    *reinterpret_cast<ulonglong2*>(w) = make_ulonglong2(pl, pc);
  }
}

// BlockLength of every block with a known start (header-parse pass).
__global__ void k_block_lengths(const uint8_t* __restrict__ src, const uint64_t* __restrict__ start,
                                const uint64_t* __restrict__ limit, uint64_t n, uint32_t* __restrict__ meta,
                                int32_t* __restrict__ status) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const uint8_t* p = src + start[b];
  const int64_t len = (int64_t)(limit[b] - start[b]);
  int64_t bl = 0;
  const int st = block_length_impl(
      len, [p](int off) { return ld_u32(p + off); }, [p](int off) { return ld_u16(p + off); }, &bl);
  const uint32_t count = len >= kHdr ? (ld_u8(p) & 0xFFu) : 0u;
  status[b] = st;
  meta[b] = st == UTL_OK ? ((uint32_t)(bl & 0xFFFFFF) | (count << 24)) : 0u;
}

__global__ void k_zero_first(uint64_t* a, uint64_t* b) {
  a[0] = 0;
  b[0] = 0;
}

__global__ void k_fill32(uint32_t* p, uint32_t v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill8(uint8_t* p, uint8_t v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace utl

using namespace utl;

size_t utl_index_tmp_bytes(uint64_t nblocks) {
  const uint64_t ntiles = (nblocks + kTile4 - 1) / kTile4;  // the smaller of the two tiles
  const uint64_t ngroups = (ntiles + kGroupTiles - 1) / kGroupTiles;
  return (size_t)(2 * (ntiles + ngroups) * sizeof(uint64_t) + 256);
}

// Passes by size: up to kInlineTiles tiles (1 Mi blocks), reduce + apply;
// above, reduce + group scan + apply, where beyond kInlineTiles groups (1 Gi
// blocks) one workgroup first scans the group totals serially.
hipError_t utl_launch_index(const uint32_t* meta, uint64_t n, uint64_t* boff, uint64_t* ooff, void* tmp,
                            hipStream_t stream) {
  if (n == 0) {
    k_zero_first<<<1, 1, 0, stream>>>(boff, ooff);
    return hipGetLastError();
  }
  if ((((uintptr_t)meta | (uintptr_t)boff | (uintptr_t)ooff) & 15) == 0) {
    const uint64_t nt = (n + kTile4 - 1) / kTile4;
    const uint64_t ng = (nt + kGroupTiles - 1) / kGroupTiles;
    uint64_t* tl = (uint64_t*)tmp;
    uint64_t* tc = tl + nt;
    uint64_t* gl = tc + nt;
    uint64_t* gc = gl + ng;
    const dim3 grid((unsigned)nt), block(kScanThreads);
    k_scan_reduce4<<<grid, block, 0, stream>>>(meta, n, tl, tc);
    if (nt <= kInlineTiles) {
      k_scan_apply4<0><<<grid, block, 0, stream>>>(meta, n, tl, tc, gl, gc, boff, ooff);
    } else {
      k_scan_groups<<<dim3((unsigned)ng), block, 0, stream>>>(tl, tc, nt, gl, gc);
      if (ng <= kInlineTiles) {
        k_scan_apply4<1><<<grid, block, 0, stream>>>(meta, n, tl, tc, gl, gc, boff, ooff);
      } else {
        k_scan_tiles<<<1, kScanThreads, 0, stream>>>(gl, gc, ng);
        k_scan_apply4<2><<<grid, block, 0, stream>>>(meta, n, tl, tc, gl, gc, boff, ooff);
      }
    }
    return hipGetLastError();
  }
  const uint64_t ntiles = (n + kTile - 1) / kTile;
  uint64_t* tl = (uint64_t*)tmp;
  uint64_t* tc = tl + ntiles;
  k_scan_reduce<<<dim3((unsigned)ntiles), dim3(kScanThreads), 0, stream>>>(meta, n, tl, tc);
  k_scan_tiles<<<1, kScanThreads, 0, stream>>>(tl, tc, ntiles);
  k_scan_apply<<<dim3((unsigned)ntiles), dim3(kScanThreads), 0, stream>>>(meta, n, tl, tc, boff, ooff);
  return hipGetLastError();
}

// Scratch of utl_index_build_noalloc: the scan's tile and group totals in a
// fixed device array (each device holds its own copy until the process
// ends), for callers that cannot allocate. Its device address is looked up
// once per device at context creation (utl_index_scratch_init); host code
// uses that address only, never the array's name.
namespace utl {
constexpr size_t kStaticScratchBytes = 1u << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];
}  // namespace utl

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

size_t utl_index_static_scratch_bytes() { return kStaticScratchBytes; }

hipError_t utl_index_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (__atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE)) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(utl::scratch));  // the current device's copy
  if (e != hipSuccess) return e;
  __atomic_store_n(&g_scratch_ptr[device], p, __ATOMIC_RELEASE);
  return hipSuccess;
}

void* utl_index_scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  return __atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE);
}

size_t utl_wave_index_entries(uint64_t nblocks) { return (size_t)(2 * ((nblocks + 7) / 8)); }

hipError_t utl_launch_wave_index(const uint32_t* meta, uint64_t n, uint64_t* wave, void* tmp, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if ((((uintptr_t)meta | (uintptr_t)wave) & 15) != 0) return hipErrorInvalidValue;
  const uint64_t nt = (n + kWaveTile - 1) / kWaveTile;  // fewer than the kTile4 tiles tmp is sized for
  const uint64_t ng = (nt + kGroupTiles - 1) / kGroupTiles;
  uint64_t* tl = (uint64_t*)tmp;
  uint64_t* tc = tl + nt;
  uint64_t* gl = tc + nt;
  uint64_t* gc = gl + ng;
  const dim3 grid((unsigned)nt), block(kScanThreads);
  k_wave_reduce<<<grid, block, 0, stream>>>(meta, n, tl, tc);
  if (nt <= kInlineTiles) {
    k_wave_apply<0><<<grid, block, 0, stream>>>(meta, n, tl, tc, gl, gc, wave);
  } else {
    k_scan_groups<<<dim3((unsigned)ng), block, 0, stream>>>(tl, tc, nt, gl, gc);
    if (ng <= kInlineTiles) {
      k_wave_apply<1><<<grid, block, 0, stream>>>(meta, n, tl, tc, gl, gc, wave);
    } else {
      k_scan_tiles<<<1, kScanThreads, 0, stream>>>(gl, gc, ng);
      k_wave_apply<2><<<grid, block, 0, stream>>>(meta, n, tl, tc, gl, gc, wave);
    }
  }
  return hipGetLastError();
}

hipError_t utl_launch_block_lengths(const uint8_t* src, const uint64_t* start, const uint64_t* limit, uint64_t n,
                                    uint32_t* meta, int32_t* status, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_block_lengths<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream>>>(src, start, limit, n, meta, status);
  return hipGetLastError();
}

hipError_t utl_launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_fill32<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream>>>(p, v, n);
  return hipGetLastError();
}

namespace utl {
__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ __launch_bounds__(256) void k_copy1(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
// A streaming kernel of a given traffic shape: reads na and writes nb
// 16-byte chunks (grid-stride, coalesced, every read consumed): the
// practical HBM rate for a kernel that moves the same bytes in the same
// directions as another one.
__global__ __launch_bounds__(256) void k_stream_rw(uint4* __restrict__ dst, uint64_t nb,
                                                   const uint4* __restrict__ src, uint64_t na) {
  const uint64_t n = na > nb ? na : nb, step = (uint64_t)gridDim.x * blockDim.x;
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    uint4 v = make_uint4((uint32_t)i, 0, 0, 0);
    if (i < na) v = src[i];
    if (i < nb)
      dst[i] = v;
    else
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9E3779B9u && nb > 0) dst[0].x = acc;  // keeps the unwritten reads live
}
}  // namespace utl

hipError_t utl_launch_stream_rw(uint8_t* dst, uint64_t write_bytes, const uint8_t* src, uint64_t read_bytes,
                                int grid, hipStream_t stream) {
  if (write_bytes == 0 && read_bytes == 0) return hipSuccess;
  k_stream_rw<<<dim3((unsigned)grid), dim3(256), 0, stream>>>((uint4*)dst, write_bytes / 16, (const uint4*)src,
                                                               read_bytes / 16);
  return hipGetLastError();
}

hipError_t utl_launch_copy(uint8_t* dst, const uint8_t* src, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t n16 = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0 ? n / 16 : 0;
  if (n16) {
    const uint64_t blocks = (n16 + 255) / 256;
    k_copy16<<<dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, stream>>>((uint4*)dst,
                                                                                          (const uint4*)src, n16);
  }
  const uint64_t done = n16 * 16, rest = n - done;
  if (rest) {  // tail, or the whole copy when a pointer is not 16-byte aligned
    const uint64_t blocks = (rest + 255) / 256;
    k_copy1<<<dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, stream>>>(dst + done, src + done,
                                                                                         rest);
  }
  return hipGetLastError();
}

hipError_t utl_launch_fill_bytes(uint8_t* p, uint8_t v, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_fill8<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream>>>(p, v, n);
  return hipGetLastError();
}
