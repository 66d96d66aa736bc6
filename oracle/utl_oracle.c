/*
 * utl_oracle.c -- CPU restatement of the Akron/utlpfor scalar block codec.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Never part of the product.
 *
 * Each function names the Go function and file:line it restates. The Go
 * reference itself cannot be built here (no Go toolchain, and its
 * StreamVByte dependency github.com/mhr3/streamvbyte v0.4.0 is not
 * vendored), so the StreamVByte "1234" codec (Lemire, Kurz, Rupp 2018:
 * 2-bit code c per value, c+1 little-endian data bytes, LSB-first codes,
 * control bytes before data) is restated from its published definition.
 *
 * Pinning status of each part is listed in tests/test_oracle_pins.py.
 */
#include "utl_oracle.h"

#include <pthread.h>
#include <stdbool.h>
#include <string.h>

enum {
  BLOCK = 128,
  LANES = 16,
  ROWS = 8,
  SUPERWORD = 64,
  HDR = 4,
  SVBLEN = 2,
  B2LEN = 2,
  FOR64 = 8,
  EXC_BMP = 16,
};

/* header.go:8-87 */
#define T_U16 (1u << 13)
#define T_U32 (2u << 13)
#define T_U64 (3u << 13)
#define H_SPECIAL (1u << 17)
#define H_RESERVED (1u << 18)
#define H_COMBINE (1u << 19)
#define H_BLOCK256 (1u << 20)
#define H_DELTA (1u << 22)
#define H_ZIGZAG (1u << 23)

static inline uint16_t ld16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static inline uint32_t ld32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static inline uint64_t ld64(const uint8_t *p) { return (uint64_t)ld32(p) | ((uint64_t)ld32(p + 4) << 32); }
static inline void st16(uint8_t *p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static inline void st32(uint8_t *p, uint32_t v) { st16(p, (uint16_t)v); st16(p + 2, (uint16_t)(v >> 16)); }
static inline void st64(uint8_t *p, uint64_t v) { st32(p, (uint32_t)v); st32(p + 4, (uint32_t)(v >> 32)); }

/* math/bits.Len32 / Len64 */
static inline int len32(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }
static inline int popc64(uint64_t x) { return __builtin_popcountll(x); }
/* Go shift semantics: a shift count >= the width yields 0. */
static inline uint32_t shl32(uint32_t x, int s) { return s >= 32 ? 0u : x << s; }
static inline uint64_t shl64(uint64_t x, int s) { return s >= 64 ? 0ull : x << s; }

/* header.go:93 utlPayloadBytes */
static inline int payload_bytes(int bw) { return ((bw + 3) >> 2) << 6; }
/* header.go:97 forBaseBytes */
static inline int for_base_bytes(int fw) { return (1 << fw) >> 1; }
/* select_bitwidth.go:10 roundUpToStep */
static inline int round_up_step(int bw) { return ((bw + 3) / 4) * 4; }
/* exceptions.go:10 excIndexSize */
static inline int exc_index_size(int exc) { return exc <= EXC_BMP ? exc : 16; }
/* svb.go:4 svbControlByteCount */
static inline int svb_ctrl_count(int n) { return (n + 3) >> 2; }
/* svb.go:9-24 svbControlBlockSizeLUT[c] = sum over the 4 codes of (code+1) */
static inline int svb_ctrl_block_size(uint8_t c) {
  return 4 + (c & 3) + ((c >> 2) & 3) + ((c >> 4) & 3) + ((c >> 6) & 3);
}

typedef struct {
  int count, bw, int_type, exc, for_width;
  bool has_exc, delta, zigzag, special, combine;
} hdr_t;

/* header.go:116 decodeHeader */
static hdr_t decode_header(uint32_t h) {
  hdr_t d;
  d.count = (int)(h & 0xFF);
  d.bw = (int)((h >> 8) & 0x1F) * 4;
  d.int_type = (int)((h >> 13) & 3);
  d.for_width = (int)((h >> 15) & 3);
  d.exc = (int)((h >> 24) & 0xFF);
  d.has_exc = d.exc > 0;
  d.delta = (h & H_DELTA) != 0;
  d.zigzag = (h & H_ZIGZAG) != 0;
  d.special = (h & H_SPECIAL) != 0;
  d.combine = (h & H_COMBINE) != 0;
  return d;
}

/* header.go:102 encodeHeader */
static uint32_t encode_header(int count, int bw, int exc, uint32_t flags) {
  uint32_t h = (uint32_t)(count & 0xFF) | ((uint32_t)((bw / 4) & 0x1F) << 8) | flags;
  h |= (uint32_t)(exc & 0xFF) << 24;
  return h;
}

/* header.go:172 validateIntType / :184 validateIntType64 */
static int validate_int_type(int t) { return (t == 2 || t == 1) ? ORC_OK : ORC_ERR_UNSUPPORTED_TYPE; }
static int validate_int_type64(int t) { return t == 3 ? ORC_OK : ORC_ERR_UNSUPPORTED_TYPE; }

/* uint64.go:70 isFor64SingleBlock */
static inline bool is_for64_single(int t, int fw, bool combine) { return t == 3 && !combine && fw == 3; }
/* uint64.go:77 forBaseBytesForBlock */
static inline int for_base_bytes_for_block(int fw, int t, bool combine) {
  return (t == 3 && !combine && fw == 3) ? FOR64 : for_base_bytes(fw);
}
/* header.go:157 payloadOffset */
static inline int payload_offset(int forb, bool has_exc, bool combine) {
  int off = HDR;
  if (has_exc) off += SVBLEN;
  off += forb;
  if (combine) off += B2LEN;
  return off;
}

/* header.go:135 Header */
int orc_header(const uint8_t *src, size_t len, int32_t out[7]) {
  memset(out, 0, 7 * sizeof(int32_t));
  if (len < HDR) return ORC_ERR_INVALID_BUFFER;
  hdr_t h = decode_header(ld32(src));
  int err = validate_int_type(h.int_type);
  if (err) return err;
  out[0] = h.count;
  out[1] = h.bw;
  out[2] = h.exc;
  out[3] = h.delta;
  out[4] = h.for_width > 0;
  out[5] = h.zigzag;
  out[6] = h.special;
  return ORC_OK;
}

/* block_length.go:8 BlockLength */
int orc_block_length(const uint8_t *src, size_t len, int64_t *out_len) {
  *out_len = 0;
  if (len < HDR) return ORC_ERR_INVALID_BUFFER;
  uint32_t header = ld32(src);
  hdr_t h = decode_header(header);
  if (h.count > BLOCK) return ORC_ERR_INVALID_BLOCK_LENGTH;
  if (h.bw > 32) return ORC_ERR_INVALID_BUFFER;
  bool combine = h.combine && h.int_type == 3;
  if (h.int_type == 3) {
    if (header & (H_RESERVED | H_BLOCK256)) return ORC_ERR_INVALID_FLAGS;
  } else if (header & (H_RESERVED | H_COMBINE | H_BLOCK256)) {
    return ORC_ERR_INVALID_FLAGS;
  }
  int forb = for_base_bytes_for_block(h.for_width, h.int_type, combine);
  int poff = payload_offset(forb, h.has_exc, combine);
  int64_t b1 = poff + (h.bw << 4);
  if (h.has_exc) {
    if (len < HDR + SVBLEN) return ORC_ERR_INVALID_BUFFER;
    int svb_len = ld16(src + HDR);
    int idx = h.exc > EXC_BMP ? 16 : h.exc;
    b1 += idx + svb_len;
  }
  if (!combine) {
    *out_len = b1;
    return ORC_OK;
  }
  if ((int64_t)len < poff) return ORC_ERR_INVALID_BUFFER;
  /* header.go:193 readBlock2Len: 2 bytes after header, svbLen, FOR base */
  int b2off = HDR + (h.has_exc ? SVBLEN : 0) + forb;
  *out_len = b1 + ld16(src + b2off);
  return ORC_OK;
}

/* ---------------------------------------------------------------- SVB --- */

/* mhr3/streamvbyte EncodeUint32 ("1234" scheme, standard Stream VByte). */
static int svb_encode(const uint32_t *v, int n, uint8_t *out) {
  int ctrl = svb_ctrl_count(n);
  memset(out, 0, (size_t)ctrl);
  int off = ctrl;
  for (int i = 0; i < n; i++) {
    uint32_t x = v[i];
    int code = x < (1u << 8) ? 0 : x < (1u << 16) ? 1 : x < (1u << 24) ? 2 : 3;
    out[i >> 2] |= (uint8_t)(code << (2 * (i & 3)));
    for (int b = 0; b <= code; b++) out[off++] = (uint8_t)(x >> (8 * b));
  }
  return off;
}

/* mhr3/streamvbyte MaxEncodedLen: control bytes + 4 data bytes per value. */
static inline int svb_max_len(int n) { return svb_ctrl_count(n) + 4 * n; }

/* mhr3/streamvbyte DecodeUint32 on a region whose control sums were
 * validated (exceptions.go:63-86): plain sequential decode. */
static void svb_decode(const uint8_t *enc, int n, uint32_t *out) {
  int off = svb_ctrl_count(n);
  for (int i = 0; i < n; i++) {
    int len = ((enc[i >> 2] >> (2 * (i & 3))) & 3) + 1;
    uint32_t x = 0;
    for (int b = 0; b < len; b++) x |= (uint32_t)enc[off + b] << (8 * b);
    off += len;
    out[i] = x;
  }
}

/* svb.go:72 svbReadValue (reads clamped to len) */
static uint32_t svb_read_value(const uint8_t *data, int64_t len, int64_t off, int bl) {
  if (off >= len) return 0;
  int64_t end = off + bl < len ? off + bl : len;
  uint32_t acc = 0;
  for (int64_t i = off; i < end; i++) acc |= (uint32_t)data[i] << ((i - off) * 8);
  return acc;
}

/* svb.go:40 svbDecodeOneInternal */
static uint32_t svb_decode_one(const uint8_t *enc, int64_t len, int count, int index) {
  if (count <= 0 || index < 0 || index >= count) return 0;
  int64_t data_off = svb_ctrl_count(count);
  int target_cb = index >> 2, slot = index & 3;
  for (int cb = 0; cb < target_cb; cb++) {
    if (cb >= len) return 0;
    data_off += svb_ctrl_block_size(enc[cb]);
  }
  if (target_cb >= len) return 0;
  uint8_t ctrl = enc[target_cb];
  for (int s = 0; s < slot; s++) data_off += ((ctrl >> (s * 2)) & 3) + 1;
  int code = (ctrl >> (slot * 2)) & 3;
  return svb_read_value(enc, len, data_off, code + 1);
}

/* --------------------------------------------------------- lane (un)pack -- */

/* scalar.go:7-59 packLanesUTLScalar / packLanePairUTL64: value (lane L,
 * row r) occupies bits [r*bw, r*bw+bw) of lane L's bit stream, whose 32-bit
 * word w lives at payload offset w*64 + L*4. Positions >= n pack as 0. */
static void pack_lanes(uint8_t *dst, const uint32_t *values, int n, int bw) {
  if (bw == 0) return;
  uint64_t mask = bw >= 32 ? 0xFFFFFFFFull : ((1ull << bw) - 1);
  for (int lane = 0; lane < LANES; lane++) {
    uint64_t acc = 0;
    int bits = 0, out = lane * 4;
    for (int v = 0; v < ROWS; v++) {
      int idx = lane + v * LANES;
      uint32_t val = idx < n ? values[idx] : 0;
      acc |= ((uint64_t)val & mask) << bits;
      bits += bw;
      while (bits >= 32) {
        st32(dst + out, (uint32_t)acc);
        out += SUPERWORD;
        acc >>= 32;
        bits -= 32;
      }
    }
    if (bits > 0) st32(dst + out, (uint32_t)acc);
  }
}

/* scalar.go:65-116 unpackLanesUTLScalar / unpackLanePairUTL64 */
static void unpack_lanes(uint32_t *dst, const uint8_t *payload, int plen, int count, int bw) {
  if (bw == 0) {
    memset(dst, 0, (size_t)count * 4);
    return;
  }
  uint32_t mask = bw >= 32 ? 0xFFFFFFFFu : ((1u << bw) - 1);
  for (int lane = 0; lane < LANES; lane++) {
    uint64_t acc = 0;
    int bits = 0, in = lane * 4;
    for (int v = 0; v < ROWS; v++) {
      while (bits < bw) {
        /* the pair loop guards an 8-byte read covering this lane */
        int pair_in = in - (lane & 1) * 4;
        if (pair_in + 8 > plen) {
          bits = bw;
          break;
        }
        acc |= (uint64_t)ld32(payload + in) << bits;
        in += SUPERWORD;
        bits += 32;
      }
      int idx = lane + v * LANES;
      if (idx < count) dst[idx] = (uint32_t)acc & mask;
      acc >>= bw;
      bits -= bw;
    }
  }
}

/* ------------------------------------------------------ FOR / bitwidth --- */

/* select_bitwidth.go:143 excCostBytesPerValue, :117 estimateExceptionCost */
static const int exc_cost_bpv[9] = {1, 1, 1, 2, 2, 3, 3, 4, 4};
static inline int estimate_exc_cost(int exc, int step_delta) {
  int bpv = exc_cost_bpv[step_delta];
  int ctrl = (exc + 3) >> 2;
  return SVBLEN + (exc < EXC_BMP ? exc : EXC_BMP) + ctrl + exc * bpv;
}

/* select_bitwidth.go:56 chooseBestFromHist */
static void choose_best_from_hist(const int hist[9], int max_si, int *best_w, int *best_exc) {
  int bw = max_si << 2, cost = HDR + (max_si << 6), exc_best = 0, above = 0;
  for (int si = max_si - 1; si >= 0; si--) {
    above += hist[si + 1];
    int c = HDR + (si << 6) + estimate_exc_cost(above, max_si - si);
    if (c < cost) {
      cost = c;
      bw = si << 2;
      exc_best = above;
    }
  }
  *best_w = bw;
  *best_exc = exc_best;
}

/* select_bitwidth.go:41 selectBitWidth */
static void select_bit_width(const uint32_t *v, int n, int *bw, int *exc) {
  int hist[9] = {0};
  uint32_t orall = 0;
  for (int i = 0; i < n; i++) {
    orall |= v[i];
    hist[(len32(v[i]) + 3) >> 2]++;
  }
  choose_best_from_hist(hist, (len32(orall) + 3) >> 2, bw, exc);
}

/* select_bitwidth.go:20 selectBitWidthNoPatch */
static int select_bit_width_nopatch(const uint32_t *v, int n) {
  uint32_t orall = 0;
  for (int i = 0; i < n; i++) orall |= v[i];
  return round_up_step(len32(orall));
}

/* for.go:11 selectFORWidth */
static int select_for_width(uint32_t minv) {
  static const int lut[5] = {0, 1, 2, 3, 3};
  return lut[(len32(minv) + 7) >> 3];
}

/* for.go:95 decideFORFull (via for.go:122 selectBitWidthWithFOR) */
static void select_bw_with_for(const uint32_t *v, int n, bool *use, uint32_t *base, int *fw) {
  uint32_t mn = v[0], mx = v[0];
  for (int i = 1; i < n; i++) {
    if (v[i] < mn) mn = v[i];
    if (v[i] > mx) mx = v[i];
  }
  *use = false;
  *base = 0;
  *fw = 0;
  if (mn == 0) return;
  int for_max = round_up_step(len32(mx - mn));
  int raw_max = round_up_step(len32(mx));
  if (for_max >= raw_max) return;
  int std_w, dummy;
  select_bit_width(v, n, &std_w, &dummy);
  if (for_max >= std_w) return;
  *use = true;
  *base = mn;
  *fw = select_for_width(mn);
}

/* for.go:17 readFORBase */
static uint32_t read_for_base(const uint8_t *buf, int off, int fw) {
  switch (fw) {
  case 1: return buf[off];
  case 2: return ld16(buf + off);
  case 3: return ld32(buf + off);
  default: return 0;
  }
}

/* for.go:32 writeFORBase */
static void write_for_base(uint8_t *buf, uint32_t base, int fw, bool has_exc) {
  int off = HDR + (has_exc ? SVBLEN : 0);
  switch (fw) {
  case 1: buf[off] = (uint8_t)base; break;
  case 2: st16(buf + off, (uint16_t)base); break;
  case 3: st32(buf + off, base); break;
  default: break;
  }
}

/* ---------------------------------------------------------------- delta -- */

static inline uint32_t zz_enc(uint32_t v) { return (v << 1) ^ (uint32_t)((int32_t)v >> 31); }
static inline uint32_t zz_dec(uint32_t v) { return (v >> 1) ^ (uint32_t)(-(int32_t)(v & 1)); }

/* delta.go:23 deltaEncodePerLaneScalar */
static bool delta_encode(uint32_t *v, int count) {
  bool zz = false;
  for (int lane = 0; lane < LANES; lane++) {
    for (int r = ROWS - 1; r > 0; r--) {
      int cur = lane + r * LANES, prev = lane + (r - 1) * LANES;
      if (cur >= count || prev >= count) continue;
      if (v[cur] < v[prev]) zz = true;
      v[cur] = v[cur] - v[prev];
    }
  }
  if (zz)
    for (int i = 0; i < count; i++) v[i] = zz_enc(v[i]);
  return zz;
}

/* delta.go:50 deltaDecodePerLaneScalar */
static void delta_decode(uint32_t *v, int count, bool zz) {
  for (int lane = 0; lane < LANES; lane++) {
    if (lane >= count) break;
    if (zz) v[lane] = zz_dec(v[lane]);
    for (int r = 1; r < ROWS; r++) {
      int cur = lane + r * LANES;
      if (cur >= count) break;
      v[cur] = v[cur - LANES] + (zz ? zz_dec(v[cur]) : v[cur]);
    }
  }
}

/* delta.go:83 deltaDecodePerLaneWithOverflowScalar (v-major first wrap) */
static int delta_decode_overflow(uint32_t *v, int count) {
  int pos = 0;
  for (int r = 1; r < ROWS; r++) {
    for (int lane = 0; lane < LANES; lane++) {
      int cur = lane + r * LANES, prev = lane + (r - 1) * LANES;
      if (cur >= count) break;
      uint32_t next = v[prev] + v[cur];
      if (pos == 0 && next < v[prev]) pos = cur;
      v[cur] = next;
    }
  }
  return pos;
}

/* ----------------------------------------------------------------- pack --- */

/* exceptions.go:138 collectAndWriteExceptions */
static void collect_and_write_exceptions(const uint32_t *v, int n, int bw, uint8_t *idx_dst,
                                         int exc, uint32_t *high) {
  if (bw >= 32) return;
  uint32_t mask = (1u << bw) - 1;
  if (exc <= EXC_BMP) {
    int k = 0;
    for (int i = 0; i < n; i++)
      if (v[i] > mask) {
        idx_dst[k] = (uint8_t)i;
        high[k] = v[i] >> bw;
        k++;
      }
  } else {
    uint64_t w0 = 0, w1 = 0;
    for (int i = 0; i < n; i++) {
      if (v[i] <= mask) continue;
      if (i < 64) w0 |= 1ull << i;
      else w1 |= 1ull << (i - 64);
    }
    st64(idx_dst, w0);
    st64(idx_dst + 8, w1);
    int k = 0;
    for (uint64_t w = w0; w; w &= w - 1) high[k++] = v[__builtin_ctzll(w)] >> bw;
    for (uint64_t w = w1; w; w &= w - 1) high[k++] = v[64 + __builtin_ctzll(w)] >> bw;
  }
}

/* scalar.go:126 packBlockScalar. Returns the block length via *out_len. */
static int pack_block(uint32_t flag, uint32_t *values, int n, uint8_t *dst, uint32_t type_flags,
                      bool combine, int *out_len) {
  if (n == 0 || n > BLOCK) return ORC_ERR_INVALID_BUFFER;
  bool no_in_place = (flag & ORC_NOINPLACE) != 0;
  uint32_t scratch[2 * BLOCK];
  uint32_t hflags = type_flags | ((flag & ORC_SPECIAL) << 14);
  uint32_t *work = values;
  bool redirected = false;

  bool use_for = false;
  uint32_t base = 0;
  int fw = 0;
  if (!(flag & ORC_NOFOR)) {
    select_bw_with_for(values, n, &use_for, &base, &fw);
    if (use_for) {
      if (no_in_place) {
        work = scratch;
        for (int i = 0; i < n; i++) work[i] = values[i] - base;
        redirected = true;
      } else {
        for (int i = 0; i < n; i++) values[i] -= base;
      }
      hflags |= (uint32_t)fw << 15;
    }
  }
  if (flag & ORC_DELTA) {
    if (no_in_place && !redirected) {
      work = scratch;
      memcpy(work, values, (size_t)n * 4);
    }
    if (delta_encode(work, n)) hflags |= H_ZIGZAG;
    hflags |= H_DELTA;
  }
  int bw, exc = 0;
  if (flag & ORC_NOPATCH) bw = select_bit_width_nopatch(work, n);
  else select_bit_width(work, n, &bw, &exc);
  int psize = payload_bytes(bw);
  int forb = for_base_bytes(fw);

  if (exc == 0) {
    int poff = payload_offset(forb, false, combine);
    memset(dst, 0, (size_t)(poff + psize)); /* combine's block2Len is patched later */
    st32(dst, encode_header(n, bw, 0, hflags));
    if (use_for) write_for_base(dst, base, fw, false);
    pack_lanes(dst + poff, work, n, bw);
    *out_len = poff + psize;
    return ORC_OK;
  }
  int idx_size = exc_index_size(exc);
  int poff = payload_offset(forb, true, combine);
  memset(dst, 0, (size_t)(poff + psize));
  st32(dst, encode_header(n, bw, exc, hflags));
  if (use_for) write_for_base(dst, base, fw, true);
  pack_lanes(dst + poff, work, n, bw);
  uint32_t high[BLOCK];
  int exc_off = poff + psize;
  collect_and_write_exceptions(work, n, bw, dst + exc_off, exc, high);
  int svb_off = exc_off + idx_size;
  int svb_len = svb_encode(high, exc, dst + svb_off);
  st16(dst + HDR, (uint16_t)svb_len);
  *out_len = svb_off + svb_len;
  return ORC_OK;
}

/* scalar.go:239 packUint32Scalar (+ pack.go:76 Append => NoInPlace) */
int orc_pack_u32(uint32_t flag, uint32_t *values, size_t n, uint8_t *dst, size_t *out_len) {
  *out_len = 0;
  flag |= (flag & ORC_APPEND) << 1;
  if (n == 0 || n > BLOCK) return ORC_ERR_INVALID_BUFFER;
  int len = 0;
  int err = pack_block(flag & ~(uint32_t)ORC_APPEND, values, (int)n, dst, T_U32, false, &len);
  if (err) return err;
  *out_len = (size_t)len;
  return ORC_OK;
}

/* --------------------------------------------------------------- unpack --- */

typedef struct {
  int consumed, poff, count, bw, exc;
  uint32_t for_base;
  bool has_exc, delta, zigzag, has_for, u64single, empty;
  int64_t block_end;
} uhdr_t;

/* unpack_block.go:62 decodeUnpackHeader */
static int decode_unpack_header(const uint8_t *buf, int64_t len, bool for_u64, uhdr_t *o) {
  memset(o, 0, sizeof(*o));
  if (len < HDR) return ORC_ERR_INVALID_BUFFER;
  hdr_t h = decode_header(ld32(buf));
  bool has_for = h.for_width > 0;
  bool combine = h.combine;
  if (for_u64) {
    int e = validate_int_type64(h.int_type);
    if (e) return e;
  } else {
    int e = validate_int_type(h.int_type);
    if (e) return e;
    combine = false;
  }
  if ((unsigned)(h.count - 1) >= BLOCK || (unsigned)h.bw > 32) {
    if (h.count == 0) {
      o->empty = true;
      return ORC_OK;
    }
    if (h.count > BLOCK) return ORC_ERR_INVALID_BLOCK_LENGTH;
    return ORC_ERR_INVALID_BUFFER;
  }
  bool u64single = for_u64 && is_for64_single(h.int_type, h.for_width, combine);
  int poff = HDR;
  if (h.has_exc) poff += SVBLEN;
  int for_off = poff;
  if (has_for) poff += u64single ? FOR64 : for_base_bytes(h.for_width);
  if (combine) poff += B2LEN;
  int pbytes = payload_bytes(h.bw);
  int idx = h.has_exc ? exc_index_size(h.exc) : 0;
  int64_t end = poff + pbytes + idx;
  if (len < end) return ORC_ERR_INVALID_BUFFER;
  if (h.has_exc) {
    int svb_len = ld16(buf + HDR);
    if (end + svb_len > len) return ORC_ERR_INVALID_BUFFER;
    end += svb_len;
  }
  o->for_base = (has_for && !u64single) ? read_for_base(buf, for_off, h.for_width) : 0;
  o->has_exc = h.has_exc;
  o->delta = h.delta;
  o->zigzag = h.zigzag;
  o->has_for = has_for;
  o->u64single = u64single;
  o->consumed = poff + pbytes;
  o->poff = poff;
  o->count = h.count;
  o->bw = h.bw;
  o->exc = h.exc;
  o->block_end = end;
  return ORC_OK;
}

/* exceptions.go:32 decodeExceptionHighBitsInto (buf is buf[:blockEnd]) */
static int decode_exc_high_bits(uint32_t *dst, const uint8_t *buf, int64_t len, int exc_start, int exc,
                                int64_t *consumed) {
  if (len < HDR + SVBLEN) return ORC_ERR_INVALID_BUFFER;
  int svb_len = ld16(buf + HDR);
  int64_t svb_start = exc_start + exc_index_size(exc);
  if (svb_start + svb_len > len) return ORC_ERR_INVALID_BUFFER;
  int ctrl = svb_ctrl_count(exc);
  if (ctrl > svb_len) return ORC_ERR_INVALID_BUFFER;
  const uint8_t *cb = buf + svb_start;
  int needed = ctrl;
  for (int g = 0; g < exc / 4; g++) needed += svb_ctrl_block_size(cb[g]);
  int k = exc % 4;
  if (k > 0) {
    uint8_t c = cb[exc / 4];
    for (int s = 0; s < k; s++) needed += ((c >> (s * 2)) & 3) + 1;
  }
  if (needed > svb_len) return ORC_ERR_INVALID_BUFFER;
  svb_decode(cb, exc, dst);
  *consumed = svb_start + svb_len;
  return ORC_OK;
}

/* exceptions.go:246 applyBitmapExceptions */
static void apply_bitmap_exceptions(uint32_t *dst, const uint32_t *dec, int ndec, const uint8_t *bmp, int count) {
  uint64_t w0 = ld64(bmp), w1 = ld64(bmp + 8);
  if (count < BLOCK) {
    if (count < 64) {
      w1 = 0;
      w0 &= (1ull << count) - 1;
    } else {
      w1 &= shl64(1ull, count - 64) - 1;
    }
  }
  if (popc64(w0) + popc64(w1) > ndec) return;
  int k = 0;
  for (; w0; w0 &= w0 - 1) dst[__builtin_ctzll(w0)] |= dec[k++];
  for (; w1; w1 &= w1 - 1) dst[64 + __builtin_ctzll(w1)] |= dec[k++];
}

/* exceptions.go:194 applyExceptions */
static int apply_exceptions(uint32_t *dst, const uint8_t *buf, int64_t len, int exc_start, int count, int bw,
                            int exc, int64_t *consumed) {
  if (exc > count) return ORC_ERR_INVALID_BUFFER;
  uint32_t dec[256];
  int err = decode_exc_high_bits(dec, buf, len, exc_start, exc, consumed);
  if (err) return err;
  if (exc <= EXC_BMP) {
    const uint8_t *pos = buf + exc_start;
    if (count == BLOCK) {
      for (int i = 0; i < exc; i++) dst[pos[i] & 0x7F] |= shl32(dec[i], bw);
    } else {
      for (int i = 0; i < exc; i++)
        if (pos[i] < count) dst[pos[i]] |= shl32(dec[i], bw);
    }
  } else {
    for (int i = 0; i < exc; i++) dec[i] = shl32(dec[i], bw);
    apply_bitmap_exceptions(dst, dec, exc, buf + exc_start, count);
  }
  return ORC_OK;
}

/* scalar.go:260 unpackUint32Scalar. dst holds 128 values. */
static int unpack_block(uint32_t *dst, const uint8_t *buf, int64_t len, bool for_u64, int *count,
                        int64_t *consumed, int64_t *ovf, uhdr_t *hout) {
  uhdr_t h;
  *count = 0;
  *consumed = 0;
  *ovf = 0;
  int err = decode_unpack_header(buf, len, for_u64, &h);
  if (hout) *hout = h;
  if (err) return err;
  if (h.empty) {
    *consumed = HDR;
    return ORC_OK;
  }
  len = h.block_end;
  unpack_lanes(dst, buf + h.poff, payload_bytes(h.bw), BLOCK, h.bw);
  int64_t cons = h.consumed;
  if (h.has_exc) {
    err = apply_exceptions(dst, buf, len, h.consumed, h.count, h.bw, h.exc, &cons);
    if (err) return err;
  }
  if (h.delta) {
    if (h.zigzag) {
      delta_decode(dst, h.count, true);
    } else {
      int p = delta_decode_overflow(dst, h.count);
      if (p > 0) {
        *ovf = p;
        return ORC_ERR_OVERFLOW;
      }
    }
  }
  if (h.has_for && !h.u64single)
    for (int i = 0; i < h.count; i++) dst[i] += h.for_base;
  *count = h.count;
  *consumed = cons;
  return ORC_OK;
}

int orc_unpack_u32(const uint8_t *src, size_t len, uint32_t *dst, int32_t *count, int64_t *consumed,
                   int64_t *ovf_pos) {
  int c = 0;
  int err = unpack_block(dst, src, (int64_t)len, false, &c, consumed, ovf_pos, NULL);
  *count = c;
  return err;
}

/* ------------------------------------------------------------------ get --- */

/* get.go:18 deltaFullUnpackThreshold at the AVX-512 dispatch level */
static inline int delta_full_unpack_threshold(int bw) { return bw >= 20 ? 4 : ROWS; }

/* get.go:34 extractPackedValueUTL */
static uint32_t extract_packed(int pos, const uint8_t *payload, int plen, int bw) {
  int lane = pos % LANES, row = pos / LANES;
  int bitpos = row * bw, word = bitpos / 32, bitoff = bitpos % 32;
  int boff = word * SUPERWORD + lane * 4;
  uint64_t acc = 0;
  if (boff + 4 <= plen) acc = ld32(payload + boff);
  if (bw > 32 - bitoff) {
    int nb = boff + SUPERWORD;
    if (nb + 4 <= plen) acc |= (uint64_t)ld32(payload + nb) << 32;
  }
  acc >>= bitoff;
  uint64_t mask = bw >= 32 ? 0xFFFFFFFFull : ((1ull << bw) - 1);
  return (uint32_t)(acc & mask);
}

/* get.go:363 bitmapRank128 */
static int bitmap_rank128(const uint8_t *bmp, int64_t blen, int byte_idx, int bit_idx) {
  if (blen < 16) return 0;
  uint64_t w0 = ld64(bmp), w1 = ld64(bmp + 8);
  if (byte_idx < 8) return popc64(w0 & ((1ull << (8 * byte_idx + bit_idx)) - 1));
  return popc64(w0) + popc64(w1 & ((1ull << (8 * (byte_idx - 8) + bit_idx)) - 1));
}

/* exceptions.go:293 findExceptionIndex. region = buf[excStart:] (len rlen) */
static int find_exception_index(const uint8_t *region, int64_t rlen, int exc, int pos) {
  if (exc <= EXC_BMP) {
    int64_t n = exc < rlen ? exc : rlen;
    for (int64_t i = 0; i < n; i++) {
      if (region[i] == (uint8_t)pos) return (int)i;
      if (region[i] > (uint8_t)pos) return -1;
    }
    return -1;
  }
  int64_t blen = rlen < 16 ? rlen : 16;
  if (blen == 0) return -1;
  int byte_idx = pos / 8, bit_idx = pos % 8;
  if (byte_idx >= blen) return -1;
  if (!(region[byte_idx] & (1 << bit_idx))) return -1;
  return bitmap_rank128(region, blen, byte_idx, bit_idx);
}

/* get.go:228 getValueDirect */
static uint32_t get_value_direct(int pos, const uint8_t *buf, int64_t len, const uint8_t *payload, int plen,
                                 int exc_start, int bw, int exc, bool has_exc) {
  uint32_t value = 0;
  if (bw > 0) value = extract_packed(pos, payload, plen, bw);
  if (has_exc) {
    const uint8_t *region = buf + exc_start;
    int64_t rlen = len - exc_start;
    int ei = find_exception_index(region, rlen, exc, pos);
    if (ei >= 0) {
      int isz = exc_index_size(exc);
      uint32_t hb = svb_decode_one(region + isz, rlen - isz, exc, ei);
      value |= shl32(hb, bw);
    }
  }
  return value;
}

/* get.go:297 applyLaneExceptions */
static void apply_lane_exceptions(uint32_t *lv, const uint8_t *region, int64_t rlen, int exc, int lane,
                                  int row, int count, int bw) {
  int isz = exc_index_size(exc);
  const uint8_t *svb = region + isz;
  int64_t svb_len = rlen - isz;
  if (exc <= EXC_BMP) {
    for (int i = 0; i < exc; i++) {
      int p = region[i];
      if (p % LANES != lane) continue;
      int v = p / LANES;
      if (v > row) break;
      lv[v] |= shl32(svb_decode_one(svb, svb_len, exc, i), bw);
    }
  } else {
    for (int v = 0; v <= row; v++) {
      int seq = lane + v * LANES;
      if (seq >= count) break;
      int byte_idx = seq / 8, bit_idx = seq % 8;
      if (!(region[byte_idx] & (1 << bit_idx))) continue;
      int rank = bitmap_rank128(region, 16, byte_idx, bit_idx);
      lv[v] |= shl32(svb_decode_one(svb, svb_len, exc, rank), bw);
    }
  }
}

/* get.go:250 getValueWithDelta. Returns ORC_REF_PANIC where Go indexes its
 * [8]uint32 lane array at row 8 (count > 128, pos in [128,144)). */
static int get_value_with_delta(int pos, const uint8_t *buf, int64_t len, const uint8_t *payload, int plen,
                                int exc_start, int bw, int count, int exc, bool has_exc, bool zz,
                                uint32_t *out) {
  int lane = pos % LANES, row = pos / LANES;
  if (row >= ROWS) return ORC_REF_PANIC;
  uint32_t lv[ROWS] = {0};
  for (int v = 0; v <= row; v++) {
    int seq = lane + v * LANES;
    if (seq >= count) break;
    if (bw > 0) lv[v] = extract_packed(seq, payload, plen, bw);
  }
  if (has_exc) apply_lane_exceptions(lv, buf + exc_start, len - exc_start, exc, lane, row, count, bw);
  if (zz) lv[0] = zz_dec(lv[0]);
  for (int v = 1; v <= row; v++) {
    int seq = lane + v * LANES;
    if (seq >= count) break;
    lv[v] = lv[v - 1] + (zz ? zz_dec(lv[v]) : lv[v]);
  }
  *out = lv[row];
  return ORC_OK;
}

/* get.go:81 getUint32Scalar */
static int get_u32_impl(int64_t pos, const uint8_t *src, int64_t len, bool for_u64, uint32_t *out) {
  *out = 0;
  if (len < HDR) return ORC_ERR_INVALID_BUFFER;
  hdr_t h = decode_header(ld32(src));
  bool has_for = h.for_width > 0;
  bool combine = h.combine;
  if (for_u64) {
    int e = validate_int_type64(h.int_type);
    if (e) return e;
  } else {
    int e = validate_int_type(h.int_type);
    if (e) return e;
    combine = false;
  }
  if (pos < 0 || pos >= h.count) return ORC_ERR_POSITION_OUT_OF_RANGE;
  if (h.bw > 32) return ORC_ERR_INVALID_BUFFER;
  bool u64single = for_u64 && is_for64_single(h.int_type, h.for_width, combine);
  int poff = HDR;
  if (h.has_exc) poff += SVBLEN;
  int forb = 0;
  if (has_for) {
    if (u64single) poff += FOR64;
    else {
      forb = for_base_bytes(h.for_width);
      poff += forb;
    }
  }
  if (combine) poff += B2LEN;
  int pbytes = payload_bytes(h.bw);
  int64_t end = poff + pbytes;
  if (len < end) return ORC_ERR_INVALID_BUFFER;
  if (h.has_exc) {
    if (h.exc > h.count) return ORC_ERR_INVALID_BUFFER;
    int svb_len = ld16(src + HDR);
    int isz = exc_index_size(h.exc);
    if (svb_len < svb_ctrl_count(h.exc) || end + isz + svb_len > len) return ORC_ERR_INVALID_BUFFER;
    end += isz + svb_len;
  }
  len = end;
  uint32_t for_base = forb ? read_for_base(src, poff - forb, h.for_width) : 0;
  const uint8_t *payload = src + poff;
  if (!h.delta) {
    *out = get_value_direct((int)pos, src, len, payload, pbytes, poff + pbytes, h.bw, h.exc, h.has_exc) +
           for_base;
    return ORC_OK;
  }
  int row = (int)(pos / LANES);
  if (row > delta_full_unpack_threshold(h.bw)) {
    /* get.go:339 getUint32FullUnpack / :198 getFullUnpackViaBlock */
    uint32_t tmp[BLOCK];
    int c;
    int64_t cons, ovf;
    int err = unpack_block(tmp, src, len, for_u64, &c, &cons, &ovf, NULL);
    if (err) return err;
    *out = tmp[pos];
    return ORC_OK;
  }
  uint32_t v;
  int err = get_value_with_delta((int)pos, src, len, payload, pbytes, poff + pbytes, h.bw, h.count, h.exc,
                                 h.has_exc, h.zigzag, &v);
  if (err) return err;
  *out = v + for_base;
  return ORC_OK;
}

int orc_get_u32(int64_t pos, const uint8_t *src, size_t len, uint32_t *out) {
  return get_u32_impl(pos, src, (int64_t)len, false, out);
}

/* ---------------------------------------------------------------- uint64 --- */

int orc_max_block_length32(uint32_t flag) {
  if (flag & ORC_NOPATCH) return HDR + ((flag & ORC_NOFOR) ? 0 : 4) + payload_bytes(32);
  return HDR + SVBLEN + ((flag & ORC_NOFOR) ? 0 : 4) + payload_bytes(28) + EXC_BMP + svb_max_len(BLOCK);
}

int orc_max_block_length64(uint32_t flag) { return 2 * orc_max_block_length32(flag) + B2LEN; }

/* uint64.go:302 packUint64Scalar (with :147 packFor64, :124
 * wrapFor64SingleBlock, :161 packUint64TwoBlock) */
int orc_pack_u64(uint32_t flag, const uint64_t *values, size_t n, uint8_t *dst, size_t *out_len) {
  *out_len = 0;
  flag |= (flag & ORC_APPEND) << 1;
  if (n == 0 || n > BLOCK) return ORC_ERR_INVALID_BUFFER;
  uint32_t inner = flag & ~(uint32_t)ORC_APPEND;
  int count = (int)n;
  uint64_t mn = values[0], mx = values[0], acc = 0;
  for (int i = 0; i < count; i++) {
    acc |= values[i];
    if (values[i] < mn) mn = values[i];
    if (values[i] > mx) mx = values[i];
  }
  uint32_t lo[BLOCK], hi[BLOCK];
  int len1 = 0, len2 = 0, err;
  if (!(inner & ORC_NOFOR) && mn > 0 && (mx - mn) < (1ull << 32)) {
    for (int i = 0; i < count; i++) lo[i] = (uint32_t)(values[i] - mn);
    uint8_t tmp[1100];
    err = pack_block(inner | ORC_NOFOR, lo, count, tmp, T_U64, false, &len1);
    if (err) return err;
    uint32_t header = ld32(tmp);
    int exc = (int)((header >> 24) & 0xFF);
    int ins = HDR + (exc > 0 ? SVBLEN : 0);
    memcpy(dst, tmp, (size_t)ins);
    st64(dst + ins, mn);
    memcpy(dst + ins + FOR64, tmp + ins, (size_t)(len1 - ins));
    st32(dst, header | (3u << 15));
    *out_len = (size_t)(len1 + FOR64);
    return ORC_OK;
  }
  if ((acc >> 32) == 0) {
    for (int i = 0; i < count; i++) lo[i] = (uint32_t)values[i];
    err = pack_block(inner, lo, count, dst, T_U64, false, &len1);
    if (err) return err;
    *out_len = (size_t)len1;
    return ORC_OK;
  }
  for (int i = 0; i < count; i++) {
    lo[i] = (uint32_t)values[i];
    hi[i] = (uint32_t)(values[i] >> 32);
  }
  err = pack_block(inner, lo, count, dst, T_U64 | H_COMBINE, true, &len1);
  if (err) return err;
  err = pack_block(inner, hi, count, dst + len1, T_U64, false, &len2);
  if (err) return err;
  hdr_t h1 = decode_header(ld32(dst));
  int b2off = HDR + (h1.has_exc ? SVBLEN : 0) + for_base_bytes(h1.for_width);
  st16(dst + b2off, (uint16_t)len2);
  *out_len = (size_t)(len1 + len2);
  return ORC_OK;
}

/* uint64.go:191 unpackUint64Block */
int orc_unpack_u64(const uint8_t *src, size_t slen, uint64_t *dst, int32_t *count, int64_t *consumed,
                   int64_t *ovf_pos) {
  int64_t len = (int64_t)slen;
  *count = 0;
  *consumed = 0;
  *ovf_pos = 0;
  if (len < HDR) return ORC_ERR_INVALID_BUFFER;
  uint32_t header = ld32(src);
  hdr_t h = decode_header(header);
  int e = validate_int_type64(h.int_type);
  if (e) return e;
  uint32_t lo[BLOCK], hi[BLOCK];
  int c1, c2;
  int64_t cons1, cons2;
  if (!h.combine) {
    uint64_t base = 0;
    if (is_for64_single(h.int_type, h.for_width, h.combine)) {
      int boff = HDR + (h.exc > 0 ? SVBLEN : 0);
      if (len < boff + FOR64) return ORC_ERR_INVALID_BUFFER;
      base = ld64(src + boff);
    }
    int err = unpack_block(lo, src, len, true, &c1, &cons1, ovf_pos, NULL);
    if (err) return err;
    for (int i = 0; i < c1; i++) dst[i] = (uint64_t)lo[i] + base;
    *count = c1;
    *consumed = cons1;
    return ORC_OK;
  }
  int err = unpack_block(lo, src, len, true, &c1, &cons1, ovf_pos, NULL);
  if (err) return err;
  int forb = for_base_bytes(h.for_width);
  int meta = payload_offset(forb, h.has_exc, true);
  if (len < meta) return ORC_ERR_INVALID_BUFFER;
  int b2len = ld16(src + HDR + (h.has_exc ? SVBLEN : 0) + forb);
  int64_t b2s = cons1;
  if (b2s + b2len > len) return ORC_ERR_INVALID_BUFFER;
  err = unpack_block(hi, src + b2s, b2len, true, &c2, &cons2, ovf_pos, NULL);
  if (err) return err;
  if (c2 != c1) return ORC_ERR_INVALID_BUFFER;
  for (int i = 0; i < c1; i++) dst[i] = ((uint64_t)hi[i] << 32) | lo[i];
  *count = c1;
  *consumed = cons1 + b2len;
  return ORC_OK;
}

/* uint64.go:372 getUint64Scalar */
int orc_get_u64(int64_t pos, const uint8_t *src, size_t slen, uint64_t *out) {
  int64_t len = (int64_t)slen;
  *out = 0;
  if (len < HDR) return ORC_ERR_INVALID_BUFFER;
  hdr_t h = decode_header(ld32(src));
  int e = validate_int_type64(h.int_type);
  if (e) return e;
  uint32_t lower, upper;
  if (!h.combine) {
    int err = get_u32_impl(pos, src, len, true, &lower);
    if (err) return err;
    uint64_t base = 0;
    if (is_for64_single(h.int_type, h.for_width, h.combine)) {
      int boff = HDR + (h.exc > 0 ? SVBLEN : 0);
      if (len < boff + FOR64) return ORC_ERR_INVALID_BUFFER;
      base = ld64(src + boff);
    }
    *out = (uint64_t)lower + base;
    return ORC_OK;
  }
  int err = get_u32_impl(pos, src, len, true, &lower);
  if (err) return err;
  int forb = for_base_bytes(h.for_width);
  int meta = payload_offset(forb, h.has_exc, true);
  if (len < meta) return ORC_ERR_INVALID_BUFFER;
  int b2len = ld16(src + HDR + (h.has_exc ? SVBLEN : 0) + forb);
  int64_t b1len = meta + (h.bw << 4);
  if (h.has_exc) {
    if (len < HDR + SVBLEN) return ORC_ERR_INVALID_BUFFER;
    int svb_len = ld16(src + HDR);
    b1len += (h.exc < EXC_BMP ? h.exc : EXC_BMP) + svb_len;
  }
  if (b1len + b2len > len) return ORC_ERR_INVALID_BUFFER;
  err = get_u32_impl(pos, src + b1len, b2len, true, &upper);
  if (err) return err;
  *out = ((uint64_t)upper << 32) | lower;
  return ORC_OK;
}

/* ------------------------------------------------------------ streams --- */

typedef struct {
  int kind; /* 0: pack32 1: pack64 2: unpack32 3: unpack64 */
  uint32_t flag;
  const void *values;
  size_t n;
  uint8_t *dst;
  const uint8_t *src;
  const uint64_t *off, *out_off;
  void *out;
  int32_t *status;
  uint32_t *block_len, *block_count;
  size_t b0, b1;
  int err;
} job_t;

static void *run_job(void *arg) {
  job_t *j = (job_t *)arg;
  for (size_t b = j->b0; b < j->b1 && !j->err; b++) {
    if (j->kind == 0 || j->kind == 1) {
      size_t lo = b * BLOCK, n = j->n - lo < BLOCK ? j->n - lo : BLOCK, len = 0;
      int err;
      if (j->kind == 0) {
        uint32_t tmp[BLOCK];
        memcpy(tmp, (const uint32_t *)j->values + lo, n * 4);
        err = orc_pack_u32(j->flag | ORC_NOINPLACE, tmp, n, j->dst + b * 2048, &len);
      } else {
        err = orc_pack_u64(j->flag, (const uint64_t *)j->values + lo, n, j->dst + b * 2048, &len);
      }
      if (err) j->err = err;
      j->block_len[b] = (uint32_t)len;
      j->block_count[b] = (uint32_t)n;
    } else {
      int32_t c;
      int64_t cons, ovf;
      int err;
      const uint8_t *p = j->src + j->off[b];
      size_t len = (size_t)(j->off[b + 1] - j->off[b]);
      uint64_t slot = j->out_off[b + 1] - j->out_off[b];
      /* a block decoding to more values than its index slot: 100 (the
       * product's UTL_ERR_ARGUMENT), nothing written */
      if (j->kind == 2) {
        uint32_t tmp[BLOCK];
        err = orc_unpack_u32(p, len, tmp, &c, &cons, &ovf);
        if (!err && (uint64_t)c > slot) err = 100;
        if (!err) memcpy((uint32_t *)j->out + j->out_off[b], tmp, (size_t)c * 4);
      } else {
        uint64_t tmp[BLOCK];
        err = orc_unpack_u64(p, len, tmp, &c, &cons, &ovf);
        if (!err && (uint64_t)c > slot) err = 100;
        if (!err) memcpy((uint64_t *)j->out + j->out_off[b], tmp, (size_t)c * 8);
      }
      j->status[b] = err | (int32_t)(err == ORC_ERR_OVERFLOW ? (ovf << 8) : 0);
    }
  }
  return NULL;
}

static int run_parallel(job_t *proto, size_t nblocks, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  job_t jobs[256];
  size_t per = (nblocks + (size_t)nthreads - 1) / (size_t)nthreads;
  int started = 0;
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = *proto;
    jobs[t].b0 = per * (size_t)t;
    jobs[t].b1 = jobs[t].b0 + per < nblocks ? jobs[t].b0 + per : nblocks;
    if (jobs[t].b0 >= jobs[t].b1) break;
    if (nthreads == 1) {
      run_job(&jobs[t]);
      started = 1;
      break;
    }
    pthread_create(&th[t], NULL, run_job, &jobs[t]);
    started = t + 1;
  }
  int err = 0;
  for (int t = 0; t < started; t++) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    if (jobs[t].err && !err) err = jobs[t].err;
  }
  return err;
}

/* Packs into 2048-byte slots, then compacts. */
static int64_t pack_stream(int kind, uint32_t flag, const void *values, size_t n, uint8_t *dst, size_t cap,
                           uint32_t *block_len, uint32_t *block_count, int nthreads, uint8_t *slots) {
  size_t nb = (n + BLOCK - 1) / BLOCK;
  job_t p;
  memset(&p, 0, sizeof p);
  p.kind = kind;
  p.flag = flag;
  p.values = values;
  p.n = n;
  p.dst = slots;
  p.block_len = block_len;
  p.block_count = block_count;
  int err = run_parallel(&p, nb, nthreads);
  if (err) return -err;
  size_t off = 0;
  for (size_t b = 0; b < nb; b++) {
    if (off + block_len[b] > cap) return -ORC_ERR_INVALID_BUFFER;
    memmove(dst + off, slots + b * 2048, block_len[b]);
    off += block_len[b];
  }
  return (int64_t)off;
}

#include <stdlib.h>
int64_t orc_pack_stream_u32(uint32_t flag, const uint32_t *values, size_t n, uint8_t *dst, size_t cap,
                            uint32_t *block_len, uint32_t *block_count, int nthreads) {
  size_t nb = (n + BLOCK - 1) / BLOCK;
  uint8_t *slots = (uint8_t *)malloc(nb * 2048 + 1);
  if (!slots) return -ORC_ERR_INVALID_BUFFER;
  int64_t r = pack_stream(0, flag, values, n, dst, cap, block_len, block_count, nthreads, slots);
  free(slots);
  return r;
}

int64_t orc_pack_stream_u64(uint32_t flag, const uint64_t *values, size_t n, uint8_t *dst, size_t cap,
                            uint32_t *block_len, uint32_t *block_count, int nthreads) {
  size_t nb = (n + BLOCK - 1) / BLOCK;
  uint8_t *slots = (uint8_t *)malloc(nb * 2048 + 1);
  if (!slots) return -ORC_ERR_INVALID_BUFFER;
  int64_t r = pack_stream(1, flag, values, n, dst, cap, block_len, block_count, nthreads, slots);
  free(slots);
  return r;
}

void orc_unpack_blocks_u32(const uint8_t *src, const uint64_t *off, const uint64_t *out_off, size_t nblocks,
                           uint32_t *out, int32_t *status, int nthreads) {
  job_t p;
  memset(&p, 0, sizeof p);
  p.kind = 2;
  p.src = src;
  p.off = off;
  p.out_off = out_off;
  p.out = out;
  p.status = status;
  run_parallel(&p, nblocks, nthreads);
}

void orc_unpack_blocks_u64(const uint8_t *src, const uint64_t *off, const uint64_t *out_off, size_t nblocks,
                           uint64_t *out, int32_t *status, int nthreads) {
  job_t p;
  memset(&p, 0, sizeof p);
  p.kind = 3;
  p.src = src;
  p.off = off;
  p.out_off = out_off;
  p.out = out;
  p.status = status;
  run_parallel(&p, nblocks, nthreads);
}

int64_t orc_walk_stream(const uint8_t *src, size_t len, uint64_t *starts, uint32_t *lens, uint32_t *counts,
                        size_t cap, int32_t *status) {
  size_t pos = 0;
  int64_t n = 0;
  *status = ORC_OK;
  while (pos < len) {
    int64_t bl = 0;
    int err = orc_block_length(src + pos, len - pos, &bl);
    if (err == ORC_OK && (uint64_t)bl > len - pos) err = ORC_ERR_INVALID_BUFFER;
    if (err != ORC_OK) {
      *status = err;
      break;
    }
    if ((size_t)n < cap) {
      starts[n] = pos;
      lens[n] = (uint32_t)bl;
      counts[n] = src[pos];
    }
    ++n;
    pos += (size_t)bl;
  }
  return n;
}
