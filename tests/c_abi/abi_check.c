/*
 * abi_check.c -- the C-ABI driven from plain C, the way a cgo binding calls
 * it: include/utlpfor_gpu.h compiled as C99 with warnings as errors, the
 * library linked as a shared object.
 *
 * Host part (no device): MaxBlockLength32/64 known answers
 * (max_block_length_test.go:12-44, uint64_test.go:14-48), Header and
 * BlockLength of the reference's golden blocks (the testdata .bin files; inputs from
 * format_conformance_test.go:693-725).
 * Device part (--gpu): PackUint32 of the golden inputs gives the golden
 * bytes, UnpackUint32 and GetUint32 give the inputs back, and PackUint64 of
 * 1e12 + {0, 1, 2} takes 19 bytes (example_test.go:206-216).
 *
 * usage: abi_check <golden dir> [--gpu]; exit status 0 when every check holds.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "utlpfor_gpu.h"

static int failures = 0;

#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);           \
      fputc('\n', stderr);                    \
      ++failures;                             \
    }                                         \
  } while (0)

static uint8_t *read_file(const char *dir, const char *name, size_t *n) {
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  uint8_t *buf = malloc(4096);
  *n = fread(buf, 1, 4096, f);
  fclose(f);
  return buf;
}

struct golden {
  const char *file;
  uint32_t flag;
  int bw, exc, delta;
  uint32_t values[128];
};

static void fill(struct golden *g) {
  for (uint32_t i = 0; i < 128; ++i) {
    if (!strcmp(g->file, "plain_128_bw8.bin")) g->values[i] = 128 + i;
    else if (!strcmp(g->file, "plain_128_bw0.bin")) g->values[i] = 0;
    else if (!strcmp(g->file, "delta_128.bin")) g->values[i] = 100 * i;
    else g->values[i] = (i % 10 == 5 && i < 100) ? 0x10000000u + i : i;
  }
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <golden dir> [--gpu]\n", argv[0]);
    return 2;
  }
  int gpu = argc > 2 && !strcmp(argv[2], "--gpu");

  /* MaxBlockLength32 {0: 1018, NoPatch: 520, NoPatch|NoFOR: 516, NoFOR: 1014}, u64 {2038, 1042, 1034, 2030} */
  const uint32_t fl[4] = {0, UTL_NOPATCH, UTL_NOPATCH | UTL_NOFOR, UTL_NOFOR};
  const int m32[4] = {1018, 520, 516, 1014}, m64[4] = {2038, 1042, 1034, 2030};
  for (int i = 0; i < 4; ++i) {
    CHECK(utl_max_block_length32(fl[i]) == m32[i], "MaxBlockLength32(%u)", fl[i]);
    CHECK(utl_max_block_length64(fl[i]) == m64[i], "MaxBlockLength64(%u)", fl[i]);
  }

  struct golden gs[4] = {{"plain_128_bw8.bin", 0, 8, 0, 0, {0}},
                         {"plain_128_bw0.bin", 0, 0, 0, 0, {0}},
                         {"delta_128.bin", UTL_DELTA, 12, 0, 1, {0}},
                         {"exceptions_128.bin", 0, 8, 10, 0, {0}}};
  utl_ctx *ctx = NULL;
  if (gpu) {
    ctx = utl_ctx_create(0);
    CHECK(ctx != NULL, "utl_ctx_create(0)");
    if (!ctx) gpu = 0;
  }
  for (int k = 0; k < 4; ++k) {
    struct golden *g = &gs[k];
    fill(g);
    size_t n = 0;
    uint8_t *blk = read_file(argv[1], g->file, &n);
    CHECK(blk != NULL, "read %s", g->file);
    if (!blk) continue;
    int32_t h[7];
    CHECK(utl_header(blk, n, h) == UTL_OK, "Header(%s)", g->file);
    CHECK(h[0] == 128 && h[1] == g->bw && h[2] == g->exc && h[3] == g->delta, "Header fields of %s: %d %d %d %d",
          g->file, h[0], h[1], h[2], h[3]);
    int64_t bl = -1;
    CHECK(utl_block_length(blk, n, &bl) == UTL_OK && bl == (int64_t)n, "BlockLength(%s) = %lld", g->file,
          (long long)bl);
    if (gpu) {
      uint32_t vals[128];
      uint8_t dst[1024];
      size_t len = 0;
      memcpy(vals, g->values, sizeof vals);
      CHECK(utl_pack_u32(ctx, g->flag | UTL_NOINPLACE, vals, 128, dst, sizeof dst, &len) == UTL_OK,
            "PackUint32(%s)", g->file);
      CHECK(len == n && !memcmp(dst, blk, n), "PackUint32(%s) bytes (%zu vs %zu)", g->file, len, n);
      uint32_t out[128];
      int32_t count = 0;
      int64_t consumed = 0, ovf = 0;
      CHECK(utl_unpack_u32(ctx, blk, n, out, &count, &consumed, &ovf) == UTL_OK, "UnpackUint32(%s)", g->file);
      CHECK(count == 128 && consumed == (int64_t)n && !memcmp(out, g->values, sizeof out),
            "UnpackUint32(%s) values", g->file);
      const int64_t pos[4] = {0, 5, 64, 127};
      for (int i = 0; i < 4; ++i) {
        uint32_t v = 0;
        CHECK(utl_get_u32(ctx, pos[i], blk, n, &v) == UTL_OK && v == g->values[pos[i]], "GetUint32(%lld, %s)",
              (long long)pos[i], g->file);
      }
      uint32_t v = 0;
      CHECK(utl_get_u32(ctx, 128, blk, n, &v) == UTL_ERR_POSITION_OUT_OF_RANGE, "GetUint32(128, %s)", g->file);
    }
    free(blk);
  }
  if (gpu) {
    const uint64_t v64[3] = {1000000000000ull, 1000000000001ull, 1000000000002ull};
    uint8_t dst[2048];
    size_t len = 0;
    CHECK(utl_pack_u64(ctx, 0, v64, 3, dst, sizeof dst, &len) == UTL_OK && len == 19, "PackUint64 -> %zu bytes",
          len);
    uint64_t out[128];
    int32_t count = 0;
    int64_t consumed = 0, ovf = 0;
    CHECK(utl_unpack_u64(ctx, dst, len, out, &count, &consumed, &ovf) == UTL_OK && count == 3 &&
              consumed == (int64_t)len && !memcmp(out, v64, sizeof v64),
          "UnpackUint64 round trip");
    utl_ctx_destroy(ctx);
  }
  printf("abi_check (%s): %d failures\n", gpu ? "host + device" : "host", failures);
  return failures ? 1 : 0;
}
