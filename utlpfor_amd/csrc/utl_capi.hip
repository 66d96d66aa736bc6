// utl_capi.hip -- the C-ABI (include/utlpfor_gpu.h) over the gfx950 kernels.
//
// Per-call entry points mirror the Go functions one to one: the block (or
// the value slice) is staged through pinned host memory into a per-context
// device buffer, one kernel launch does the work, results come back with one
// copy, and the call synchronizes on the context's stream. Batched entry
// points work on device-resident data and are asynchronous.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/utlpfor_gpu.h"
#include "utl_encode.h"
#include "utl_format.h"
#include "utl_get.h"
#include "utl_kernels.h"

namespace {

// Bytes of a block that any decode or Get can touch: a uint64 two-block's
// second sub-block starts at most 16 + 512 + 16 + 65535 bytes in (svbLen is
// a u16) and reads at most ~1.7 KB after that. Copying this window of a
// longer caller buffer is exact: validation still uses the true length.
constexpr size_t kReadWindow = 72 * 1024;
constexpr size_t kStageBytes = 128 * 1024;
constexpr size_t kResultOff = 96 * 1024;  // results region inside the staging buffers
// Launch variant of the per-call decode (one block): the smallest LDS arena.
const utl_tune kOneBlockTune{2560, 4, 1, 1};

thread_local std::string g_last_error;

int hip_fail(hipError_t e, const char* what) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  g_last_error = buf;
  return UTL_ERR_HIP;
}

#define HIP_TRY(expr)                                      \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(e_, #expr);      \
  } while (0)

}  // namespace

struct utl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // a second stream for work that runs beside the context's (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  uint8_t* h_stage = nullptr;  // pinned, host-mapped, fine-grained (coherent)
  uint8_t* d_stage = nullptr;  // the device's address of h_stage
  void* d_tmp = nullptr;
  size_t d_tmp_bytes = 0;
  unsigned long long* d_counter = nullptr;
  utl_tune tune{2560, 4, 0, 4};  // small arena, 4 waves per workgroup (the tuner may change it)
  int cu_count = 0;
  char name[128] = {0};
};

namespace {

int ctx_init(utl_ctx* c, int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
    g_last_error = "no HIP device";
    return UTL_ERR_NO_DEVICE;
  }
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_last_error = std::string("device is not gfx950: ") + prop.gcnArchName;
    return UTL_ERR_NO_DEVICE;
  }
  c->device = device;
  c->cu_count = prop.multiProcessorCount;
  snprintf(c->name, sizeof c->name, "%s (%s)", prop.name, prop.gcnArchName);
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  // Per-call staging: kernels read a call's inputs from and write its
  // results to this host buffer directly (fine-grained, so no GPU cache
  // holds a stale line between calls): one launch and one synchronization
  // per call, no copies.
  HIP_TRY(hipHostMalloc((void**)&c->h_stage, kStageBytes, hipHostMallocMapped | hipHostMallocCoherent));
  HIP_TRY(hipHostGetDevicePointer((void**)&c->d_stage, c->h_stage, 0));
  HIP_TRY(hipMalloc((void**)&c->d_counter, 4096));  // 8 per-XCD work counters, 64 B apart
  HIP_TRY(utl_index_scratch_init(device));             // the fixed scratch of utl_index_build_noalloc
  c->tune.grid_wgs = c->cu_count * 5;  // five persistent u32 workgroups are resident per CU
  return UTL_OK;
}

void ctx_free(utl_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->d_tmp) hipFree(c->d_tmp);
  if (c->d_counter) hipFree(c->d_counter);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->side) hipStreamSynchronize(c->side);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

// The per-thread default context lives until the process ends: freeing HIP
// objects from a thread-exit destructor would run after the HIP runtime (or
// a profiler layered on it) may already have torn down its own thread state.
struct TlsCtx {
  utl_ctx* c = nullptr;
};
thread_local TlsCtx g_tls;

// The context to use: the caller's, or this thread's default on device 0.
utl_ctx* resolve(utl_ctx* c, int* err) {
  *err = UTL_OK;
  if (!c) {
    if (!g_tls.c) {
      utl_ctx* n = new utl_ctx();
      *err = ctx_init(n, 0);
      if (*err) {
        ctx_free(n);
        return nullptr;
      }
      g_tls.c = n;
    }
    c = g_tls.c;
  }
  const hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) {
    *err = hip_fail(e, "hipSetDevice");
    return nullptr;
  }
  return c;
}

int ensure_tmp(utl_ctx* c, size_t bytes) {
  if (c->d_tmp_bytes >= bytes) return UTL_OK;
  if (c->d_tmp) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(c->d_tmp));
    c->d_tmp = nullptr;
    c->d_tmp_bytes = 0;
  }
  size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
  HIP_TRY(hipMalloc(&c->d_tmp, want));
  c->d_tmp_bytes = want;
  return UTL_OK;
}

// Per-call unpack/get staging layout (both buffers):
//   [0,16)  block_off {0, len}   [16,32) out_off {0, 128}
//   [32,48) request pos / block  [64, 64 + window) block bytes
//   kResultOff: status i32, consumed i32, then 128 values of 8 bytes
int stage_block(utl_ctx* c, const uint8_t* src, size_t len) {
  const size_t ncopy = len < kReadWindow ? len : kReadWindow;
  uint64_t* hdr = (uint64_t*)c->h_stage;
  hdr[0] = 0;
  hdr[1] = len;
  hdr[2] = 0;
  hdr[3] = 128;
  if (ncopy) memcpy(c->h_stage + 64, src, ncopy);
  return UTL_OK;
}

}  // namespace

extern "C" {

utl_ctx* utl_ctx_create(int device) {
  utl_ctx* c = new utl_ctx();
  if (ctx_init(c, device) != UTL_OK) {
    ctx_free(c);
    return nullptr;
  }
  return c;
}

void utl_ctx_destroy(utl_ctx* ctx) { ctx_free(ctx); }

void* utl_ctx_stream(utl_ctx* ctx) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  return c ? (void*)c->stream : nullptr;
}

int utl_ctx_sync(utl_ctx* ctx) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UTL_OK;
}

int utl_ctx_device_info(utl_ctx* ctx, char* name, size_t name_cap, int* cu_count) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if (name && name_cap) snprintf(name, name_cap, "%s", c->name);
  if (cu_count) *cu_count = c->cu_count;
  return UTL_OK;
}

const char* utl_last_error(void) { return g_last_error.c_str(); }

// ---- header-only entry points (host reads of at most 6 bytes) ----

int utl_max_block_length32(uint32_t flag) {
  // scalar.go:336-357: NoPatch => header + FOR + 32-bit payload; else the
  // bw=28 payload with all 128 values as exceptions in bitmap mode.
  if (flag & UTL_NOPATCH) return utl::kHdr + ((flag & UTL_NOFOR) ? 0 : 4) + utl::payload_bytes(32);
  return utl::kHdr + utl::kSvbLen + ((flag & UTL_NOFOR) ? 0 : 4) + utl::payload_bytes(28) + utl::kExcBitmap +
         (utl::svb_ctrl_count(utl::kBlock) + 4 * utl::kBlock);
}

int utl_max_block_length64(uint32_t flag) { return 2 * utl_max_block_length32(flag) + utl::kB2Len; }

int utl_header(const uint8_t* src, size_t len, int32_t out[7]) {
  if (!out) return UTL_ERR_ARGUMENT;
  memset(out, 0, 7 * sizeof(int32_t));
  if (len < (size_t)utl::kHdr || !src) return UTL_ERR_INVALID_BUFFER;
  const utl::Hdr h = utl::decode_header(utl::host_ld32(src));
  if (h.int_type != 2 && h.int_type != 1) return UTL_ERR_UNSUPPORTED_TYPE;  // header.go:172
  out[0] = h.count;
  out[1] = h.bw;
  out[2] = h.exc;
  out[3] = h.delta;
  out[4] = h.for_width > 0;
  out[5] = h.zigzag;
  out[6] = h.special;
  return UTL_OK;
}

int utl_block_length(const uint8_t* src, size_t len, int64_t* out_len) {
  if (!out_len) return UTL_ERR_ARGUMENT;
  if (!src) len = 0;
  return utl::block_length_impl(
      (int64_t)len, [src](int off) { return utl::host_ld32(src + off); },
      [src](int off) { return utl::host_ld16(src + off); }, out_len);
}

// ---- per-call decode ----

static int unpack_call(utl_ctx* ctx, bool u64, const uint8_t* src, size_t len, void* dst, int32_t* count,
                       int64_t* consumed, int64_t* ovf_pos) {
  if (count) *count = 0;
  if (consumed) *consumed = 0;
  if (ovf_pos) *ovf_pos = 0;
  if (!dst || (!src && len)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if ((err = stage_block(c, src, len))) return err;
  uint8_t* d = c->d_stage;
  int32_t* d_status = (int32_t*)(d + kResultOff);
  int32_t* d_cons = d_status + 1;
  void* d_vals = d + kResultOff + 64;
  HIP_TRY(utl_launch_decode(u64, d + 64, (const uint64_t*)d, (const uint64_t*)(d + 16), 1, d_vals, d_status, d_cons,
                            kOneBlockTune, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int32_t st = *(volatile int32_t*)(c->h_stage + kResultOff);
  const int code = UTL_STATUS_CODE(st);
  if (code == UTL_OK) {
    // a decoded block holds its header count (0 for an empty block)
    const int out_n = (int)src[0];
    const int32_t cons = *(int32_t*)(c->h_stage + kResultOff + 4);
    memcpy(dst, c->h_stage + kResultOff + 64, (size_t)out_n * (u64 ? 8 : 4));
    if (count) *count = out_n;
    if (consumed) *consumed = cons;
  } else if (code == UTL_ERR_OVERFLOW && ovf_pos) {
    *ovf_pos = UTL_STATUS_POS(st);
  }
  return code;
}

int utl_unpack_u32(utl_ctx* ctx, const uint8_t* src, size_t len, uint32_t* dst, int32_t* count, int64_t* consumed,
                   int64_t* ovf_pos) {
  return unpack_call(ctx, false, src, len, dst, count, consumed, ovf_pos);
}

int utl_unpack_u64(utl_ctx* ctx, const uint8_t* src, size_t len, uint64_t* dst, int32_t* count, int64_t* consumed,
                   int64_t* ovf_pos) {
  return unpack_call(ctx, true, src, len, dst, count, consumed, ovf_pos);
}

// ---- per-call Get ----

static int get_call(utl_ctx* ctx, bool u64, int64_t pos, const uint8_t* src, size_t len, void* out) {
  if (!out || (!src && len)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  const size_t ncopy = len < kReadWindow ? len : kReadWindow;
  uint64_t* hdr = (uint64_t*)c->h_stage;
  hdr[0] = 0;
  hdr[1] = len;
  hdr[2] = 0;
  hdr[3] = 128;
  hdr[4] = 0;                // request block
  hdr[5] = (uint64_t)pos;    // request position
  if (ncopy) memcpy(c->h_stage + 64, src, ncopy);
  uint8_t* d = c->d_stage;
  int32_t* d_status = (int32_t*)(d + kResultOff);
  void* d_val = d + kResultOff + 64;
  HIP_TRY(utl_launch_get(u64, d + 64, (const uint64_t*)d, (const uint64_t*)(d + 16), 1, nullptr, 0,
                         (const uint64_t*)(d + 32), (const int64_t*)(d + 40), 1, d_val, d_status, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int32_t st = *(volatile int32_t*)(c->h_stage + kResultOff);
  if (u64) *(uint64_t*)out = st == UTL_OK ? *(uint64_t*)(c->h_stage + kResultOff + 64) : 0;
  else *(uint32_t*)out = st == UTL_OK ? *(uint32_t*)(c->h_stage + kResultOff + 64) : 0;
  // An overflow inside Get's full-unpack fallback is still ErrOverflow.
  return UTL_STATUS_CODE(st);
}

int utl_get_u32(utl_ctx* ctx, int64_t pos, const uint8_t* src, size_t len, uint32_t* out) {
  return get_call(ctx, false, pos, src, len, out);
}

int utl_get_u64(utl_ctx* ctx, int64_t pos, const uint8_t* src, size_t len, uint64_t* out) {
  return get_call(ctx, true, pos, src, len, out);
}

// ---- per-call pack ----

static int pack_call(utl_ctx* ctx, bool u64, uint32_t flag, const void* values, size_t n, uint32_t* inplace,
                     uint8_t* dst, size_t dst_cap, size_t* out_len) {
  if (out_len) *out_len = 0;
  flag |= (flag & UTL_APPEND) << 1;  // pack.go:79 Append implies NoInPlace
  if (n == 0 || n > (size_t)utl::kBlock) return UTL_ERR_INVALID_BUFFER;
  if (!values || !dst) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  // staging: [0,8) in_off, [8,12) in_meta, [64, 64+1024) values,
  // kResultOff: out_meta u32, then work values (u32 x 128) at +64, then the
  // block slot at +1024.
  uint64_t* hdr = (uint64_t*)c->h_stage;
  hdr[0] = 0;
  *(uint32_t*)(c->h_stage + 8) = (uint32_t)n | ((flag & 0xFF) << 8);
  memcpy(c->h_stage + 64, values, n * (u64 ? 8 : 4));
  uint8_t* d = c->d_stage;
  uint32_t* d_meta = (uint32_t*)(d + kResultOff);
  uint32_t* d_work = (uint32_t*)(d + kResultOff + 64);
  uint8_t* d_slot = d + kResultOff + 1024;
  if (u64)
    HIP_TRY(utl_launch_encode_u64((const uint64_t*)(d + 64), (const uint64_t*)d, (const uint32_t*)(d + 8), 1, d_slot,
                                  2048, d_meta, c->stream));
  else
    HIP_TRY(utl_launch_encode_u32((const uint32_t*)(d + 64), (const uint64_t*)d, (const uint32_t*)(d + 8), 1, d_slot,
                                  1024, d_meta, d_work, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const uint32_t meta = *(volatile uint32_t*)(c->h_stage + kResultOff);
  if (meta == 0xFFFFFFFFu) return UTL_ERR_INVALID_BUFFER;
  const size_t len = meta & 0xFFFFFF;
  if (out_len) *out_len = len;
  if (len > dst_cap) return UTL_ERR_ARGUMENT;
  // the block starts after the zero bytes (0..3) the encoder puts before it
  // in the slot; its first byte, the count, is not 0
  const uint8_t* blk = c->h_stage + kResultOff + 1024;
  int sh = 0;
  while (sh < 3 && blk[sh] == 0) ++sh;
  memcpy(dst, blk + sh, len);
  if (inplace && !(flag & UTL_NOINPLACE)) memcpy(inplace, c->h_stage + kResultOff + 64, n * 4);
  return UTL_OK;
}

int utl_pack_u32(utl_ctx* ctx, uint32_t flag, uint32_t* values, size_t n, uint8_t* dst, size_t dst_cap,
                 size_t* out_len) {
  return pack_call(ctx, false, flag, values, n, values, dst, dst_cap, out_len);
}

int utl_pack_u64(utl_ctx* ctx, uint32_t flag, const uint64_t* values, size_t n, uint8_t* dst, size_t dst_cap,
                 size_t* out_len) {
  return pack_call(ctx, true, flag, values, n, nullptr, dst, dst_cap, out_len);
}

// ---- device memory ----

int utl_dev_alloc(void** ptr, size_t bytes) {
  if (!ptr) return UTL_ERR_ARGUMENT;
  *ptr = nullptr;
  int err;
  if (!resolve(nullptr, &err)) return err;
  HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return UTL_OK;
}

int utl_dev_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return UTL_OK;
}

int utl_host_alloc(void** ptr, size_t bytes) {
  if (!ptr) return UTL_ERR_ARGUMENT;
  *ptr = nullptr;
  int err;
  if (!resolve(nullptr, &err)) return err;
  HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return UTL_OK;
}

int utl_host_free(void* ptr) {
  if (ptr) HIP_TRY(hipHostFree(ptr));
  return UTL_OK;
}

int utl_memcpy_h2d(utl_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UTL_OK;
}

int utl_memcpy_d2h(utl_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UTL_OK;
}

int utl_memset_d(utl_ctx* ctx, void* dst, int value, size_t bytes) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_fill_bytes((uint8_t*)dst, (uint8_t)value, bytes, c->stream));
  return UTL_OK;
}

int utl_dev_mem_info(size_t* free_bytes, size_t* total_bytes) {
  int err;
  if (!resolve(nullptr, &err)) return err;
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return UTL_OK;
}

// ---- batched, device-resident ----

int utl_index_build(utl_ctx* ctx, const uint32_t* d_meta, size_t nblocks, uint64_t* d_block_off, uint64_t* d_out_off) {
  if ((!d_meta && nblocks) || !d_block_off || !d_out_off) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if ((err = ensure_tmp(c, utl_index_tmp_bytes(nblocks)))) return err;
  HIP_TRY(utl_launch_index(d_meta, nblocks, d_block_off, d_out_off, c->d_tmp, c->stream));
  return UTL_OK;
}

int utl_block_lengths(utl_ctx* ctx, const uint8_t* d_src, const uint64_t* d_block_start, const uint64_t* d_limit,
                      size_t nblocks, uint32_t* d_meta, int32_t* d_status) {
  if (nblocks && (!d_src || !d_block_start || !d_limit || !d_meta || !d_status)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_block_lengths(d_src, d_block_start, d_limit, nblocks, d_meta, d_status, c->stream));
  return UTL_OK;
}

int utl_decode_u32(utl_ctx* ctx, const uint8_t* d_src, const uint64_t* d_block_off, const uint64_t* d_out_off,
                   size_t nblocks, uint32_t* d_out, int32_t* d_status, int32_t* d_consumed) {
  if (nblocks && (!d_src || !d_block_off || !d_out_off || !d_out || !d_status)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_decode(false, d_src, d_block_off, d_out_off, nblocks, d_out, d_status, d_consumed, c->tune,
                            c->stream));
  return UTL_OK;
}

int utl_decode_u64(utl_ctx* ctx, const uint8_t* d_src, const uint64_t* d_block_off, const uint64_t* d_out_off,
                   size_t nblocks, uint64_t* d_out, int32_t* d_status, int32_t* d_consumed) {
  if (nblocks && (!d_src || !d_block_off || !d_out_off || !d_out || !d_status)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_decode(true, d_src, d_block_off, d_out_off, nblocks, d_out, d_status, d_consumed, c->tune,
                            c->stream));
  return UTL_OK;
}

int utl_decode_persistent(utl_ctx* ctx, int is_u64, const uint8_t* d_src, const uint64_t* d_block_off,
                          const uint64_t* d_out_off, size_t nblocks, void* d_out, int32_t* d_status,
                          int32_t* d_consumed) {
  if (nblocks && (!d_src || !d_block_off || !d_out_off || !d_out || !d_status)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_decode_persistent(is_u64 != 0, d_src, d_block_off, d_out_off, nblocks, d_out, d_status,
                                       d_consumed, c->d_counter, c->tune, c->stream));
  return UTL_OK;
}

int utl_index_build_noalloc(utl_ctx* ctx, const uint32_t* d_meta, size_t nblocks, uint64_t* d_block_off,
                            uint64_t* d_out_off) {
  if (!ctx) {
    g_last_error = "utl_index_build_noalloc: needs a context from utl_ctx_create (creating the default one allocates)";
    return UTL_ERR_ARGUMENT;
  }
  if ((!d_meta && nblocks) || !d_block_off || !d_out_off) return UTL_ERR_ARGUMENT;
  const size_t need = utl_index_tmp_bytes(nblocks);
  if (need > utl_index_static_scratch_bytes()) {
    char buf[160];
    snprintf(buf, sizeof buf, "utl_index_build_noalloc: %zu blocks need %zu scratch bytes, more than the fixed %zu",
             nblocks, need, utl_index_static_scratch_bytes());
    g_last_error = buf;
    return UTL_ERR_ARGUMENT;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  void* tmp = utl_index_scratch_ptr(ctx->device);
  if (!tmp) {
    g_last_error = "utl_index_build_noalloc: the fixed scratch of this device was not set up";
    return UTL_ERR_HIP;
  }
  HIP_TRY(utl_launch_index(d_meta, nblocks, d_block_off, d_out_off, tmp, ctx->stream));
  return UTL_OK;
}

int utl_wave_index_build(utl_ctx* ctx, const uint32_t* d_meta, size_t nblocks, uint64_t* d_wave) {
  if (nblocks && (!d_meta || !d_wave)) return UTL_ERR_ARGUMENT;
  if ((((uintptr_t)d_meta | (uintptr_t)d_wave) & 15) != 0) {
    g_last_error = "utl_wave_index_build: d_meta and d_wave must be 16-byte aligned";
    return UTL_ERR_ARGUMENT;
  }
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if ((err = ensure_tmp(c, utl_index_tmp_bytes(nblocks)))) return err;
  HIP_TRY(utl_launch_wave_index(d_meta, nblocks, d_wave, c->d_tmp, c->stream));
  return UTL_OK;
}

int utl_decode_meta(utl_ctx* ctx, int is_u64, int persistent, const uint8_t* d_src, const uint32_t* d_meta,
                    const uint64_t* d_wave, size_t nblocks, void* d_out, int32_t* d_status, int32_t* d_consumed) {
  if (nblocks && (!d_src || !d_meta || !d_wave || !d_out || !d_status)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_decode_meta(is_u64 != 0, d_src, d_meta, d_wave, nblocks, d_out, d_status, d_consumed,
                                 persistent ? c->d_counter : nullptr, c->tune, c->stream));
  return UTL_OK;
}

int utl_set_tune(utl_ctx* ctx, const utl_tune_params* t) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if (!t) return UTL_ERR_ARGUMENT;
  const utl_tune n{t->arena_bytes, t->waves_per_wg, t->grid_wgs, (uint32_t)(t->tiles_per_grab > 0 ? t->tiles_per_grab : 0)};
  if (!utl_tune_valid(n)) {
    g_last_error = "utl_set_tune: arena_bytes must be 2560, 3328 or 4352, waves_per_wg 2 or 4, grid_wgs and "
                   "tiles_per_grab >= 1";
    return UTL_ERR_ARGUMENT;
  }
  c->tune = n;
  return UTL_OK;
}

int utl_get_tune(utl_ctx* ctx, utl_tune_params* t) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if (!t) return UTL_ERR_ARGUMENT;
  t->arena_bytes = c->tune.arena_bytes;
  t->waves_per_wg = c->tune.waves_per_wg;
  t->grid_wgs = c->tune.grid_wgs;
  t->tiles_per_grab = (int)c->tune.tiles_per_grab;
  return UTL_OK;
}

int utl_index_walk(utl_ctx* ctx, const uint8_t* d_src, const uint64_t* d_stream_off, size_t nstreams,
                   const uint64_t* d_blk_off, uint32_t* d_meta, uint64_t* d_block_start, uint32_t* d_nblocks,
                   int32_t* d_status) {
  if (nstreams && (!d_src || !d_stream_off || !d_nblocks || !d_status || (d_meta && !d_blk_off) ||
                   (d_block_start && !d_meta)))
    return UTL_ERR_ARGUMENT;
  if (((uintptr_t)d_src & 15) != 0) {  // the walk stages 16-byte chunks
    g_last_error = "utl_index_walk: d_src must be 16-byte aligned";
    return UTL_ERR_ARGUMENT;
  }
  if (nstreams > 0xFFFFFFFFull) {  // stream numbers are 32-bit words on the device
    g_last_error = "utl_index_walk: more than 2^32 - 1 streams";
    return UTL_ERR_ARGUMENT;
  }
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if (nstreams == 0) return UTL_OK;
  // the first and last stream offsets size the chunk records
  HIP_TRY(hipMemcpyAsync(c->h_stage, d_stream_off, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(c->h_stage + 8, d_stream_off + nstreams, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  uint64_t ends[2];
  memcpy(ends, c->h_stage, 16);
  const uint64_t total = ends[1] > ends[0] ? ends[1] - ends[0] : 0;
  if ((err = ensure_tmp(c, utl_walk_tmp_bytes(nstreams, total)))) return err;
  HIP_TRY(utl_launch_index_walk(d_src, d_stream_off, nstreams, total, d_blk_off, d_meta, d_block_start, d_nblocks,
                                d_status, c->d_tmp, c->stream, c->side, c->ev_fork, c->ev_join));
  return UTL_OK;
}

int utl_gather_index(utl_ctx* ctx, const uint64_t* d_block_off, const uint64_t* d_out_off, size_t store_nblocks,
                     const uint64_t* d_req, size_t nreq, size_t nentries, uint64_t* d_gblock_off,
                     uint64_t* d_gout_off) {
  if (nentries && (!d_block_off || !d_out_off || !d_req || !nreq || !d_gblock_off || !d_gout_off))
    return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_gather_index(d_block_off, d_out_off, store_nblocks, d_req, nreq, nentries, d_gblock_off,
                                  d_gout_off, c->stream));
  return UTL_OK;
}

int utl_get_batch(utl_ctx* ctx, int is_u64, const uint8_t* d_src, const uint64_t* d_block_off,
                  const uint64_t* d_out_off, size_t nblocks, const uint64_t* d_req_block, const int64_t* d_req_pos,
                  size_t nreq, void* d_out, int32_t* d_status) {
  if (nreq && (!d_src || !d_block_off || !d_out_off || !d_req_pos || !d_out || !d_status)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_get(is_u64 != 0, d_src, d_block_off, d_out_off, nblocks, nullptr, 0, d_req_block, d_req_pos,
                         nreq, d_out, d_status, c->stream));
  return UTL_OK;
}

int utl_pos_dir_build(utl_ctx* ctx, const uint64_t* d_block_off, const uint64_t* d_out_off, size_t nblocks,
                      uint64_t* d_dir, size_t ndir) {
  if (ndir && (!d_block_off || !d_out_off || !d_dir || nblocks == 0)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_pos_dir(d_block_off, d_out_off, nblocks, d_dir, ndir, c->stream));
  return UTL_OK;
}

int utl_get_batch_dir(utl_ctx* ctx, int is_u64, const uint8_t* d_src, const uint64_t* d_block_off,
                      const uint64_t* d_out_off, size_t nblocks, const uint64_t* d_dir, size_t ndir,
                      const int64_t* d_req_pos, size_t nreq, void* d_out, int32_t* d_status) {
  if (nreq && (!d_src || !d_block_off || !d_out_off || !d_dir || !ndir || !d_req_pos || !d_out || !d_status))
    return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_get(is_u64 != 0, d_src, d_block_off, d_out_off, nblocks, d_dir, ndir, nullptr, d_req_pos, nreq,
                         d_out, d_status, c->stream));
  return UTL_OK;
}

int utl_encode_blocks(utl_ctx* ctx, int is_u64, const void* d_values, const uint64_t* d_in_off,
                      const uint32_t* d_in_meta, size_t nblocks, uint8_t* d_slots, uint32_t* d_out_meta) {
  if (nblocks && (!d_values || !d_in_off || !d_in_meta || !d_slots || !d_out_meta)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if (is_u64)
    HIP_TRY(utl_launch_encode_u64((const uint64_t*)d_values, d_in_off, d_in_meta, nblocks, d_slots, UTL_SLOT_BYTES64,
                                  d_out_meta, c->stream));
  else
    HIP_TRY(utl_launch_encode_u32((const uint32_t*)d_values, d_in_off, d_in_meta, nblocks, d_slots, UTL_SLOT_BYTES32,
                                  d_out_meta, nullptr, c->stream));
  return UTL_OK;
}

int utl_encode_u32_direct(utl_ctx* ctx, const uint32_t* d_values, const uint64_t* d_in_off, const uint32_t* d_in_meta,
                          size_t nblocks, uint8_t* d_dst, size_t dst_cap, uint32_t* d_out_meta, uint64_t* d_block_off,
                          size_t* out_nbytes) {
  if (out_nbytes) *out_nbytes = 0;
  if (!d_block_off || (nblocks && (!d_values || !d_in_off || !d_in_meta || !d_dst || !d_out_meta)))
    return UTL_ERR_ARGUMENT;
  if (nblocks && dst_cap / (size_t)utl_max_block_length32(0) < nblocks) {
    g_last_error = "utl_encode_u32_direct: dst_cap must be at least nblocks * utl_max_block_length32(0)";
    return UTL_ERR_ARGUMENT;
  }
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  if (nblocks == 0) {
    HIP_TRY(utl_launch_fill_bytes((uint8_t*)d_block_off, 0, sizeof(uint64_t), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return UTL_OK;
  }
  if ((err = ensure_tmp(c, utl_encode_direct_tmp_bytes(nblocks)))) return err;
  uint64_t* lb = (uint64_t*)c->d_tmp;
  HIP_TRY(utl_launch_encode_u32_direct(d_values, d_in_off, d_in_meta, nblocks, d_dst, dst_cap, d_out_meta, d_block_off,
                                       nullptr, lb, c->stream));
  // the look-back's error word and the stream's length
  const uint64_t ntiles = (nblocks + 31) / 32;
  HIP_TRY(hipMemcpyAsync(c->h_stage, lb + 1 + ntiles, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(c->h_stage + 8, d_block_off + nblocks, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  uint64_t w[2];
  memcpy(w, c->h_stage, 16);
  if (w[0]) {
    g_last_error = (w[0] & 1) ? "utl_encode_u32_direct: a tile's look-back for its stream offset did not complete"
                              : "utl_encode_u32_direct: a block would have ended past dst_cap";
    return UTL_ERR_HIP;
  }
  if (out_nbytes) *out_nbytes = (size_t)w[1];
  return UTL_OK;
}

int utl_encode_u32_size(utl_ctx* ctx, const uint32_t* d_values, const uint64_t* d_in_off, const uint32_t* d_in_meta,
                        size_t nblocks, uint32_t* d_out_meta, uint64_t* d_plan) {
  if (nblocks && (!d_values || !d_in_off || !d_in_meta || !d_out_meta || !d_plan)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_encode_u32_size(d_values, d_in_off, d_in_meta, nblocks, d_out_meta, d_plan, c->stream));
  return UTL_OK;
}

int utl_encode_u32_write(utl_ctx* ctx, const uint32_t* d_values, const uint64_t* d_in_off, const uint32_t* d_in_meta,
                         size_t nblocks, const uint64_t* d_plan, const uint64_t* d_block_off, uint8_t* d_dst,
                         size_t dst_cap) {
  if (nblocks && (!d_values || !d_in_off || !d_in_meta || !d_plan || !d_block_off || !d_dst)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_encode_u32_write(d_values, d_in_off, d_in_meta, nblocks, d_plan, d_block_off, d_dst, dst_cap,
                                      c->stream));
  return UTL_OK;
}

int utl_encode_u64_size(utl_ctx* ctx, const uint64_t* d_values, const uint64_t* d_in_off, const uint32_t* d_in_meta,
                        size_t nblocks, uint32_t* d_out_meta) {
  if (nblocks && (!d_values || !d_in_off || !d_in_meta || !d_out_meta)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_encode_u64_size(d_values, d_in_off, d_in_meta, nblocks, d_out_meta, c->stream));
  return UTL_OK;
}

int utl_encode_u64_write(utl_ctx* ctx, const uint64_t* d_values, const uint64_t* d_in_off, const uint32_t* d_in_meta,
                         size_t nblocks, const uint64_t* d_block_off, uint8_t* d_dst, size_t dst_cap) {
  if (nblocks && (!d_values || !d_in_off || !d_in_meta || !d_block_off || !d_dst)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_encode_u64_write(d_values, d_in_off, d_in_meta, nblocks, d_block_off, d_dst, dst_cap, c->stream));
  return UTL_OK;
}

int utl_compact(utl_ctx* ctx, int is_u64, const uint8_t* d_slots, const uint64_t* d_block_off, size_t nblocks,
                uint8_t* d_dst) {
  if (nblocks && (!d_slots || !d_block_off || !d_dst)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_compact(d_slots, is_u64 ? UTL_SLOT_BYTES64 : UTL_SLOT_BYTES32, !is_u64, d_block_off, nblocks,
                             d_dst, c->stream));
  return UTL_OK;
}

int utl_make_jobs(utl_ctx* ctx, const uint64_t* d_val_off, const uint64_t* d_blk_off, const uint8_t* d_flags,
                  size_t nstreams, uint64_t* d_in_off, uint32_t* d_in_meta) {
  if (nstreams && (!d_val_off || !d_blk_off || !d_flags || !d_in_off || !d_in_meta)) return UTL_ERR_ARGUMENT;
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_make_jobs(d_val_off, d_blk_off, d_flags, nstreams, d_in_off, d_in_meta, c->stream));
  return UTL_OK;
}

// ---- HIP graph of index scan + decode ----

}  // extern "C"

// A graph owns the scratch its captured index scan writes (tile totals), so
// later calls on the same context that grow or free the context's own
// scratch never touch memory a live graph replays into.
struct utl_graph {
  utl_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* d_tmp = nullptr;
  uint64_t* d_wave = nullptr;  // the wave index, when the caller asks for no block index
};

extern "C" {

utl_graph* utl_graph_decode_create(utl_ctx* ctx, int is_u64, int persistent, const uint8_t* d_src,
                                   const uint32_t* d_meta, size_t nblocks, uint64_t* d_block_off,
                                   uint64_t* d_out_off, void* d_out, int32_t* d_status) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  const bool wave_form = !d_block_off && !d_out_off;
  if (!c || !d_meta || (!wave_form && (!d_block_off || !d_out_off)) || (nblocks && (!d_src || !d_out || !d_status)))
    return nullptr;
  if (wave_form && ((uintptr_t)d_meta & 15) != 0) {
    g_last_error = "utl_graph_decode_create: d_meta must be 16-byte aligned without a block index";
    return nullptr;
  }
  // everything the captured kernels touch must exist before capture
  if (hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
  utl_graph* g = new utl_graph();
  g->ctx = c;
  bool ok = hipMalloc(&g->d_tmp, utl_index_tmp_bytes(nblocks)) == hipSuccess;
  if (ok && wave_form)
    ok = hipMalloc((void**)&g->d_wave, (utl_wave_index_entries(nblocks) + 2) * sizeof(uint64_t)) == hipSuccess;
  if (!ok) {
    g_last_error = "graph scratch allocation failed";
    hipFree(g->d_tmp);
    delete g;
    return nullptr;
  }
  unsigned long long* counter = persistent ? c->d_counter : nullptr;
  ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    if (wave_form) {
      ok = utl_launch_wave_index(d_meta, nblocks, g->d_wave, g->d_tmp, c->stream) == hipSuccess &&
           utl_launch_decode_meta(is_u64 != 0, d_src, d_meta, g->d_wave, nblocks, d_out, d_status, nullptr, counter,
                                  c->tune, c->stream) == hipSuccess;
    } else {
      ok = utl_launch_index(d_meta, nblocks, d_block_off, d_out_off, g->d_tmp, c->stream) == hipSuccess;
      if (ok && persistent)
        ok = utl_launch_decode_persistent(is_u64 != 0, d_src, d_block_off, d_out_off, nblocks, d_out, d_status,
                                          nullptr, counter, c->tune, c->stream) == hipSuccess;
      else if (ok)
        ok = utl_launch_decode(is_u64 != 0, d_src, d_block_off, d_out_off, nblocks, d_out, d_status, nullptr,
                               c->tune, c->stream) == hipSuccess;
    }
    hipGraph_t graph = nullptr;
    const bool ended = hipStreamEndCapture(c->stream, &graph) == hipSuccess;
    ok = ok && ended && graph;
    g->graph = graph;
  }
  if (ok) ok = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) == hipSuccess;
  if (!ok) {
    g_last_error = "graph capture failed";
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    hipFree(g->d_tmp);
    hipFree(g->d_wave);
    delete g;
    return nullptr;
  }
  return g;
}

int utl_graph_launch(utl_graph* g) {
  if (!g) return UTL_ERR_ARGUMENT;
  HIP_TRY(hipSetDevice(g->ctx->device));
  HIP_TRY(hipGraphLaunch(g->exec, g->ctx->stream));
  return UTL_OK;
}

void utl_graph_destroy(utl_graph* g) {
  if (!g) return;
  hipSetDevice(g->ctx->device);
  hipStreamSynchronize(g->ctx->stream);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->d_tmp) hipFree(g->d_tmp);
  if (g->d_wave) hipFree(g->d_wave);
  delete g;
}

// ---- timing and calibration helpers ----

int utl_event_create(void** ev) {
  if (!ev) return UTL_ERR_ARGUMENT;
  int err;
  if (!resolve(nullptr, &err)) return err;
  hipEvent_t e;
  HIP_TRY(hipEventCreate(&e));
  *ev = (void*)e;
  return UTL_OK;
}

int utl_event_destroy(void* ev) {
  if (ev) HIP_TRY(hipEventDestroy((hipEvent_t)ev));
  return UTL_OK;
}

int utl_event_record(utl_ctx* ctx, void* ev) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(hipEventRecord((hipEvent_t)ev, c->stream));
  return UTL_OK;
}

int utl_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!ms) return UTL_ERR_ARGUMENT;
  HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
  HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return UTL_OK;
}

int utl_dev_copy(utl_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_copy((uint8_t*)d_dst, (const uint8_t*)d_src, bytes, c->stream));
  return UTL_OK;
}

int utl_dev_stream(utl_ctx* ctx, void* d_dst, size_t write_bytes, const void* d_src, size_t read_bytes,
                   int grid_wgs) {
  if ((write_bytes && !d_dst) || (read_bytes && !d_src) || grid_wgs < 0) return UTL_ERR_ARGUMENT;
  if ((((uintptr_t)d_dst | (uintptr_t)d_src) & 15) != 0) {
    g_last_error = "utl_dev_stream: pointers must be 16-byte aligned";
    return UTL_ERR_ARGUMENT;
  }
  int err;
  utl_ctx* c = resolve(ctx, &err);
  if (!c) return err;
  HIP_TRY(utl_launch_stream_rw((uint8_t*)d_dst, write_bytes, (const uint8_t*)d_src, read_bytes,
                               grid_wgs ? grid_wgs : 4 * c->cu_count, c->stream));
  return UTL_OK;
}

}  // extern "C"
