/*
 * utl_oracle.h -- CPU restatement of the Akron/utlpfor scalar block codec.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the GPU product is
 * compared against; it is never linked into, loaded by, or called from the
 * product library. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every function follows the Go reference literally (file:line cited at each
 * definition in utl_oracle.c): error precedence, Go shift semantics
 * (x << 32 == 0), clamped StreamVByte reads, the AVX-512 Get threshold.
 */
#ifndef UTL_ORACLE_H
#define UTL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes; numerically identical to the product's UTL_* codes. */
enum {
  ORC_OK = 0,
  ORC_ERR_INVALID_BUFFER = 1,        /* utlpfor.go:71 ErrInvalidBuffer */
  ORC_ERR_POSITION_OUT_OF_RANGE = 2, /* utlpfor.go:74 */
  ORC_ERR_INVALID_BLOCK_LENGTH = 3,  /* utlpfor.go:77 */
  ORC_ERR_INVALID_FLAGS = 4,         /* utlpfor.go:80 */
  ORC_ERR_UNSUPPORTED_TYPE = 5,      /* utlpfor.go:86 */
  ORC_ERR_OVERFLOW = 6,              /* utlpfor.go:87 *ErrOverflow{Position} */
  /* The Go reference panics (index out of range) on this input. Only the
   * oracle reports this; the product returns ORC_ERR_INVALID_BUFFER. */
  ORC_REF_PANIC = 99,
};

/* Flag bits, utlpfor.go:21-68 */
enum {
  ORC_DELTA = 1,
  ORC_NOFOR = 2,
  ORC_NOPATCH = 4,
  ORC_SPECIAL = 8,
  ORC_APPEND = 16,
  ORC_NOINPLACE = 32,
};

/* header.go:135 Header. out[7] = count, bitWidth, excCount, hasDelta, hasFOR,
 * hasZigZag, hasSpecial. */
int orc_header(const uint8_t *src, size_t len, int32_t out[7]);

/* block_length.go:8 BlockLength */
int orc_block_length(const uint8_t *src, size_t len, int64_t *out_len);

/* scalar.go:336 MaxBlockLength32, uint64.go:451 MaxBlockLength64 */
int orc_max_block_length32(uint32_t flag);
int orc_max_block_length64(uint32_t flag);

/* pack.go:76 PackUint32 -> packUint32Scalar. Writes the block at dst (capacity
 * >= 1018). values is mutated in place exactly as the reference does unless
 * NoInPlace or Append is set. Append has no positional meaning here: the
 * caller concatenates. */
int orc_pack_u32(uint32_t flag, uint32_t *values, size_t n, uint8_t *dst,
                 size_t *out_len);

/* unpack.go:5 UnpackUint32 -> unpackUint32Scalar. dst must hold 128 values.
 * On ORC_ERR_OVERFLOW, *ovf_pos holds ErrOverflow.Position. */
int orc_unpack_u32(const uint8_t *src, size_t len, uint32_t *dst,
                   int32_t *count, int64_t *consumed, int64_t *ovf_pos);

/* get.go:74 GetUint32 (AVX-512 dispatch level for the delta threshold). */
int orc_get_u32(int64_t pos, const uint8_t *src, size_t len, uint32_t *out);

/* uint64.go:279 PackUint64 (dst capacity >= 2038; values never mutated). */
int orc_pack_u64(uint32_t flag, const uint64_t *values, size_t n, uint8_t *dst,
                 size_t *out_len);

/* uint64.go:338 UnpackUint64. dst must hold 128 values. */
int orc_unpack_u64(const uint8_t *src, size_t len, uint64_t *dst,
                   int32_t *count, int64_t *consumed, int64_t *ovf_pos);

/* uint64.go:364 GetUint64 */
int orc_get_u64(int64_t pos, const uint8_t *src, size_t len, uint64_t *out);

/* ---- stream helpers (the caller pattern of example_test.go:171-187) ---- */

/* Pack n values as consecutive 128-value blocks with flag|Append semantics
 * (values untouched). Writes per-block byte lengths and value counts.
 * Returns total bytes, or -(error code). */
int64_t orc_pack_stream_u32(uint32_t flag, const uint32_t *values, size_t n,
                            uint8_t *dst, size_t cap, uint32_t *block_len,
                            uint32_t *block_count, int nthreads);
int64_t orc_pack_stream_u64(uint32_t flag, const uint64_t *values, size_t n,
                            uint8_t *dst, size_t cap, uint32_t *block_len,
                            uint32_t *block_count, int nthreads);

/* Decode nblocks blocks: block k is UnpackUint32(src[off[k]:off[k+1]]) and
 * its values go to out[out_off[k]...]. status[k] = code | (ovf_pos << 8).
 * Uses nthreads POSIX threads over contiguous block ranges. */
void orc_unpack_blocks_u32(const uint8_t *src, const uint64_t *off,
                           const uint64_t *out_off, size_t nblocks,
                           uint32_t *out, int32_t *status, int nthreads);
void orc_unpack_blocks_u64(const uint8_t *src, const uint64_t *off,
                           const uint64_t *out_off, size_t nblocks,
                           uint64_t *out, int32_t *status, int nthreads);

/* The caller's serial walk of one stream (example_test.go:171-187): block
 * k+1 starts BlockLength(block k) bytes after block k. Stops at the end of
 * src (*status = 0), at the first BlockLength error (*status = that error)
 * or at a block longer than the bytes left (*status = ErrInvalidBuffer).
 * Writes the first min(blocks, cap) blocks' start offsets, byte lengths and
 * value counts (header byte 0); returns the number of blocks walked. */
int64_t orc_walk_stream(const uint8_t *src, size_t len, uint64_t *starts, uint32_t *lens, uint32_t *counts,
                        size_t cap, int32_t *status);

#ifdef __cplusplus
}
#endif
#endif
