// utl_container.hip -- the stream container of include/utlpfor_gpu.h (host
// code only): blocks exactly as the reference packs them, per-block meta
// words and a stream directory, with every structural check made at parse
// time so that a loaded container can go straight to the device index scan.
#include <stdint.h>
#include <string.h>

#include "../../include/utlpfor_gpu.h"

static_assert(sizeof(utl_container_header) == 80, "container header layout");
static_assert(sizeof(utl_stream_entry) == 48, "stream entry layout");

namespace {

constexpr uint64_t kHeader = sizeof(utl_container_header);
constexpr uint64_t kEntry = sizeof(utl_stream_entry);

bool mul_ok(uint64_t a, uint64_t b, uint64_t* r) { return !__builtin_mul_overflow(a, b, r); }
bool add_ok(uint64_t a, uint64_t b, uint64_t* r) { return !__builtin_add_overflow(a, b, r); }

// Section offsets for the counts; false on overflow.
bool layout(uint64_t nstreams, uint64_t nblocks, uint64_t nbytes, uint64_t* dir_off, uint64_t* meta_off,
            uint64_t* data_off, uint64_t* total) {
  uint64_t dir_bytes, meta_bytes;
  if (!mul_ok(nstreams, kEntry, &dir_bytes) || !mul_ok(nblocks, 4, &meta_bytes)) return false;
  if (!add_ok(meta_bytes, 7, &meta_bytes)) return false;
  meta_bytes &= ~7ull;
  *dir_off = kHeader;
  return add_ok(*dir_off, dir_bytes, meta_off) && add_ok(*meta_off, meta_bytes, data_off) &&
         add_ok(*data_off, nbytes, total);
}

// The directory tiles the blocks, values and bytes in order, and each
// stream's meta words sum to its bytes and values.
bool check_dir(const utl_stream_entry* dir, uint64_t nstreams, const uint32_t* meta, uint64_t nblocks,
               uint64_t nbytes, uint64_t* nvalues) {
  uint64_t b = 0, v = 0, o = 0;
  for (uint64_t s = 0; s < nstreams; ++s) {
    const utl_stream_entry& e = dir[s];
    if (e.first_block != b || e.first_value != v || e.byte_off != o || e.is_u64 > 1) return false;
    if (e.nblocks > nblocks - b) return false;
    uint64_t sb = 0, sv = 0;
    for (uint64_t k = b; k < b + e.nblocks; ++k) {
      const uint32_t len = meta[k] & 0xFFFFFFu, cnt = meta[k] >> 24;
      if (len < 4 || cnt == 0 || cnt > 128) return false;
      sb += len;
      sv += cnt;
    }
    if (sv != e.nvalues || sb > nbytes - o) return false;
    b += e.nblocks;
    v += sv;
    o += sb;
  }
  *nvalues = v;
  return b == nblocks && o == nbytes;
}

}  // namespace

extern "C" {

size_t utl_container_bytes(size_t nstreams, size_t nblocks, size_t nbytes) {
  uint64_t d, m, x, total;
  return layout(nstreams, nblocks, nbytes, &d, &m, &x, &total) ? (size_t)total : 0;
}

int utl_container_write(uint8_t* dst, size_t dst_cap, const utl_stream_entry* dir, size_t nstreams,
                        const uint32_t* meta, size_t nblocks, const uint8_t* data, size_t nbytes,
                        size_t* out_len) {
  if (!out_len) return UTL_ERR_ARGUMENT;
  *out_len = 0;
  if ((nstreams && !dir) || (nblocks && !meta) || (nbytes && !data)) return UTL_ERR_ARGUMENT;
  uint64_t dir_off, meta_off, data_off, total, nvalues = 0;
  if (!layout(nstreams, nblocks, nbytes, &dir_off, &meta_off, &data_off, &total)) return UTL_ERR_ARGUMENT;
  if (!check_dir(dir, nstreams, meta, nblocks, nbytes, &nvalues)) return UTL_ERR_ARGUMENT;
  if (!dst || dst_cap < total) {
    *out_len = (size_t)total;  // the size needed
    return UTL_ERR_ARGUMENT;
  }
  utl_container_header h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, UTL_CONTAINER_MAGIC, 8);
  h.version = UTL_CONTAINER_VERSION;
  h.header_bytes = (uint32_t)kHeader;
  h.nstreams = nstreams;
  h.nblocks = nblocks;
  h.nbytes = nbytes;
  h.nvalues = nvalues;
  h.dir_off = dir_off;
  h.meta_off = meta_off;
  h.data_off = data_off;
  h.total_bytes = total;
  memcpy(dst, &h, kHeader);
  if (nstreams) memcpy(dst + dir_off, dir, nstreams * kEntry);
  if (nblocks) memcpy(dst + meta_off, meta, nblocks * 4);
  memset(dst + meta_off + nblocks * 4, 0, data_off - meta_off - nblocks * 4);
  if (nbytes) memcpy(dst + data_off, data, nbytes);
  *out_len = (size_t)total;
  return UTL_OK;
}

int utl_container_parse(const uint8_t* src, size_t len, utl_container_header* hdr, const utl_stream_entry** dir,
                        const uint32_t** meta, const uint8_t** data) {
  if (!hdr || !dir || !meta || !data) return UTL_ERR_ARGUMENT;
  if (src && (reinterpret_cast<uintptr_t>(src) & 7) != 0) return UTL_ERR_ARGUMENT;  // 8-byte sections
  memset(hdr, 0, sizeof *hdr);
  *dir = nullptr;
  *meta = nullptr;
  *data = nullptr;
  if (!src || len < kHeader) return UTL_ERR_INVALID_BUFFER;
  utl_container_header h;
  memcpy(&h, src, kHeader);
  if (memcmp(h.magic, UTL_CONTAINER_MAGIC, 8) != 0 || h.version != UTL_CONTAINER_VERSION ||
      h.header_bytes != kHeader)
    return UTL_ERR_INVALID_BUFFER;
  uint64_t dir_off, meta_off, data_off, total, nvalues = 0;
  if (!layout(h.nstreams, h.nblocks, h.nbytes, &dir_off, &meta_off, &data_off, &total)) return UTL_ERR_INVALID_BUFFER;
  if (h.dir_off != dir_off || h.meta_off != meta_off || h.data_off != data_off || h.total_bytes != total ||
      total > len)
    return UTL_ERR_INVALID_BUFFER;
  const utl_stream_entry* d = reinterpret_cast<const utl_stream_entry*>(src + dir_off);
  const uint32_t* m = reinterpret_cast<const uint32_t*>(src + meta_off);
  if (!check_dir(d, h.nstreams, m, h.nblocks, h.nbytes, &nvalues) || nvalues != h.nvalues)
    return UTL_ERR_INVALID_BUFFER;
  *hdr = h;
  *dir = d;
  *meta = m;
  *data = src + data_off;
  return UTL_OK;
}

}  // extern "C"
