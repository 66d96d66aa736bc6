/*
 * utlpfor_gpu.h -- C-ABI of the MI355X-native UTL-PFOR block codec.
 *
 * Drop-in boundary for the Akron/utlpfor Go API (reference file:line given at
 * each entry point). Plain pointers and sizes only; every compute entry point
 * runs hand-written HIP kernels for gfx950.
 *
 * Error codes map 1:1 to the Go sentinels (utlpfor.go:71-115). ErrOverflow's
 * Position is returned through an out-parameter. Codes >= 100 are failures of
 * this library or of the HIP runtime and have no Go counterpart.
 *
 * Threading: package-level entry points take a utl_ctx*; passing NULL uses a
 * per-thread default context, so they are reentrant like the Go functions.
 * A utl_ctx (one HIP stream + staging memory) corresponds to a Go
 * Compressor32/64 and, like it, is not safe for concurrent use
 * (compressor.go:7-8).
 */
#ifndef UTLPFOR_GPU_H
#define UTLPFOR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (utlpfor.go:71-115) ---- */
#define UTL_OK 0
#define UTL_ERR_INVALID_BUFFER 1        /* ErrInvalidBuffer */
#define UTL_ERR_POSITION_OUT_OF_RANGE 2 /* ErrPositionOutOfRange */
#define UTL_ERR_INVALID_BLOCK_LENGTH 3  /* ErrInvalidBlockLength */
#define UTL_ERR_INVALID_FLAGS 4         /* ErrInvalidFlags */
#define UTL_ERR_UNSUPPORTED_TYPE 5      /* ErrUnsupportedType */
#define UTL_ERR_OVERFLOW 6              /* *ErrOverflow{Position} */
#define UTL_ERR_ARGUMENT 100            /* bad C-ABI argument (null pointer, short capacity) */
#define UTL_ERR_HIP 101                 /* HIP runtime failure */
#define UTL_ERR_NO_DEVICE 102           /* no usable gfx950 device */

/* Per-block status words written by the batched entry points:
 * status = code | (overflow_position << 8). */
#define UTL_STATUS_CODE(s) ((s) & 0xFF)
#define UTL_STATUS_POS(s) ((s) >> 8)

/* ---- flags (utlpfor.go:21-68) ---- */
#define UTL_DELTA 1u
#define UTL_NOFOR 2u
#define UTL_NOPATCH 4u
#define UTL_SPECIAL 8u
#define UTL_APPEND 16u
#define UTL_NOINPLACE 32u

/* pack.go:5,11 and uint64.go:15 */
#define UTL_SCRATCH_LEN 128
#define UTL_SCRATCH_LEN_NOINPLACE 256
#define UTL_SCRATCH_LEN64 384

typedef struct utl_ctx utl_ctx;

/* ---- context (compressor.go:16-107 NewUint32/NewUint64) ---- */
/* Creates a context on `device` (HIP ordinal): one HIP stream, pinned host
 * staging and device scratch. Returns NULL if no device is usable. */
utl_ctx *utl_ctx_create(int device);
void utl_ctx_destroy(utl_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own
 * work against it. */
void *utl_ctx_stream(utl_ctx *ctx);
/* Blocks until all work queued on the context's stream has finished. */
int utl_ctx_sync(utl_ctx *ctx);
/* Human-readable device name and compute-unit count of the context. */
int utl_ctx_device_info(utl_ctx *ctx, char *name, size_t name_cap, int *cu_count);

/* ---- per-call drop-in: one reference call = one C-ABI call ---- */

/* scalar.go:336 MaxBlockLength32 */
int utl_max_block_length32(uint32_t flag);
/* uint64.go:451 MaxBlockLength64 */
int utl_max_block_length64(uint32_t flag);

/* header.go:135 Header. out[7] = count, bitWidth, excCount, hasDelta, hasFOR,
 * hasZigZag, hasSpecial. A 4-byte host read: no device work. */
int utl_header(const uint8_t *src, size_t len, int32_t out[7]);

/* block_length.go:8 BlockLength. A <=6-byte host read: no device work; the
 * device form for whole streams is utl_block_lengths. */
int utl_block_length(const uint8_t *src, size_t len, int64_t *out_len);

/* pack.go:76 PackUint32. Encodes values[0:n] (1 <= n <= 128) on the GPU into
 * dst (capacity dst_cap >= utl_max_block_length32(flag) recommended; a shorter
 * capacity that cannot hold the block returns UTL_ERR_ARGUMENT with the
 * needed size in *out_len). Append has no positional meaning here: the caller
 * concatenates. Like the reference, values is overwritten with the post-FOR,
 * post-delta, post-zigzag work values unless NoInPlace or Append is set. */
int utl_pack_u32(utl_ctx *ctx, uint32_t flag, uint32_t *values, size_t n, uint8_t *dst, size_t dst_cap,
                 size_t *out_len);

/* unpack.go:5 UnpackUint32. Decodes one block on the GPU into dst, which
 * must hold 128 values or at least the block's count (src[0]): only the
 * *count decoded values are written. *consumed as the reference returns it.
 * On UTL_ERR_OVERFLOW, *ovf_pos = ErrOverflow.Position. */
int utl_unpack_u32(utl_ctx *ctx, const uint8_t *src, size_t len, uint32_t *dst, int32_t *count,
                   int64_t *consumed, int64_t *ovf_pos);

/* get.go:74 GetUint32 (delta full-unpack threshold of the reference's
 * AVX-512 dispatch level, get.go:18-28). */
int utl_get_u32(utl_ctx *ctx, int64_t pos, const uint8_t *src, size_t len, uint32_t *out);

/* uint64.go:279 PackUint64 (values are never modified). */
int utl_pack_u64(utl_ctx *ctx, uint32_t flag, const uint64_t *values, size_t n, uint8_t *dst,
                 size_t dst_cap, size_t *out_len);

/* uint64.go:338 UnpackUint64. dst must hold 128 values or at least the
 * block's count (src[0]); only the *count decoded values are written. */
int utl_unpack_u64(utl_ctx *ctx, const uint8_t *src, size_t len, uint64_t *dst, int32_t *count,
                   int64_t *consumed, int64_t *ovf_pos);

/* uint64.go:364 GetUint64 */
int utl_get_u64(utl_ctx *ctx, int64_t pos, const uint8_t *src, size_t len, uint64_t *out);

/* ---- device memory (the batched forms work on device-resident data) ---- */
int utl_dev_alloc(void **ptr, size_t bytes);
int utl_dev_free(void *ptr);
/* Pinned (page-locked) host memory: copies to and from it run at the full
 * link rate. No counterpart in the reference (Go owns host memory); used by
 * the serving loop to deliver results. */
int utl_host_alloc(void **ptr, size_t bytes);
int utl_host_free(void *ptr);
int utl_memcpy_h2d(utl_ctx *ctx, void *dst, const void *src, size_t bytes);  /* synchronous */
int utl_memcpy_d2h(utl_ctx *ctx, void *dst, const void *src, size_t bytes);  /* synchronous */
int utl_memset_d(utl_ctx *ctx, void *dst, int value, size_t bytes);          /* fill kernel, async */
int utl_dev_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- batched, device-resident forms (no reference counterpart; each block
 * result equals the per-call result on the same bytes) ----
 *
 * A stream index describes nblocks concatenated blocks: block_off[k] is the
 * byte offset of block k in src and block_off[k+1] its end, so block k is
 * decoded exactly as UnpackUint32(src[block_off[k]:block_off[k+1]]), which on
 * every block a serial BlockLength walk reaches equals the walk's result
 * (block_length.go:8, example_test.go:171-187). out_off[k] is the value
 * offset of block k in the decoded output. All pointers prefixed d_ are
 * device pointers. Calls are asynchronous on the context's stream. */

/* Device exclusive scan that replaces the serial BlockLength walk: from
 * per-block meta words (byte length | value count << 24, as the encoders
 * emit them) build block_off[nblocks+1] and out_off[nblocks+1]. */
int utl_index_build(utl_ctx *ctx, const uint32_t *d_meta, size_t nblocks, uint64_t *d_block_off,
                    uint64_t *d_out_off);

/* utl_index_build that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). No counterpart in the reference. The scan's
 * scratch is a fixed 1 MiB device array per GPU (set up by utl_ctx_create),
 * so nblocks may be at most 67,026,944 (UTL_ERR_ARGUMENT beyond, with the
 * reason in utl_last_error); ctx must not be NULL (creating the default
 * context allocates). The scratch is shared by every call on the device:
 * calls that overlap on one GPU (other contexts or streams) race. Output is
 * identical to utl_index_build's. */
int utl_index_build_noalloc(utl_ctx *ctx, const uint32_t *d_meta, size_t nblocks, uint64_t *d_block_off,
                            uint64_t *d_out_off);

/* Header-parse pass over a device stream whose block starts are known:
 * BlockLength (block_length.go:8) of block k = src[start[k]:limit[k]], as
 * meta words (length | count << 24) and per-block status. */
int utl_block_lengths(utl_ctx *ctx, const uint8_t *d_src, const uint64_t *d_block_start, const uint64_t *d_limit,
                      size_t nblocks, uint32_t *d_meta, int32_t *d_status);

/* Index of foreign streams whose block starts are unknown: the reference's
 * serial walk (block k+1 starts BlockLength(block k) bytes after block k,
 * block_length.go:8, example_test.go:171-187) on the device, a lane per
 * short stream and chunk-parallel for long ones (an entry vote per chunk,
 * verified against the previous chunk's exit, so the result is the serial
 * walk's for any input). Stream s is d_src[d_stream_off[s] :
 * d_stream_off[s+1]]. At most 2^32 - 1 streams. Its
 * walk stops at the stream's end (d_status[s] = 0), at the first BlockLength
 * error (that error), or at a block longer than the bytes left
 * (ErrInvalidBuffer); d_nblocks[s] counts the blocks before the stop. With
 * d_meta NULL only counts and statuses are written (pass 1); otherwise the
 * blocks' meta words (length | count << 24, as the encoders emit them) go to
 * d_meta[d_blk_off[s] ..] and, when d_block_start is not NULL, their byte
 * offsets in d_src to d_block_start[d_blk_off[s] ..], d_blk_off being the
 * exclusive scan of the pass-1 counts (pass 2). Block k then decodes as
 * d_src[start[k] : start[k+1]]: its exact bytes inside a stream, the rest of
 * its stream after a stop (as UnpackUint32(buf[pos:]) reads it,
 * example_test.go:183). utl_index_build over d_meta gives the value offsets.
 * d_src must be 16-byte aligned and d_stream_off non-decreasing; no byte
 * of a stream past its end is read. The call reads d_stream_off[0] and
 * d_stream_off[nstreams] back to size its scratch (one synchronization of
 * the context's stream); the rest is asynchronous. */
int utl_index_walk(utl_ctx *ctx, const uint8_t *d_src, const uint64_t *d_stream_off, size_t nstreams,
                   const uint64_t *d_blk_off, uint32_t *d_meta, uint64_t *d_block_start, uint32_t *d_nblocks,
                   int32_t *d_status);

/* Wave index of a stream described by its meta words: d_wave[2w] and
 * d_wave[2w+1] are the byte and value offsets of block 8w, for w <
 * ceil(nblocks / 8) (the same device prefix scan of block lengths as
 * utl_index_build, kept at one entry per eight blocks: 2 bytes per block
 * instead of 16). d_meta and d_wave must be 16-byte aligned. */
int utl_wave_index_build(utl_ctx *ctx, const uint32_t *d_meta, size_t nblocks, uint64_t *d_wave);

/* Batched UnpackUint32 / UnpackUint64 (is_u64) of concatenated blocks in
 * order, found from their meta words and the wave index (each wavefront of
 * the decode scans its eight meta words itself): results identical to
 * utl_decode_u32/u64 over the index utl_index_build makes from the same meta
 * words. persistent selects the work-queue form. */
int utl_decode_meta(utl_ctx *ctx, int is_u64, int persistent, const uint8_t *d_src, const uint32_t *d_meta,
                    const uint64_t *d_wave, size_t nblocks, void *d_out, int32_t *d_status, int32_t *d_consumed);

/* Batched UnpackUint32 over an index (one launch). d_status[k] receives the
 * block's status word; d_consumed (optional, may be NULL) its consumed
 * bytes. Block k's values go to d_out[d_out_off[k] ..]; values of failed
 * blocks are not written, nor is any slot position past a block's count.
 * The index may jump between regions of d_src (a gathered index). d_out is
 * naturally aligned. */
int utl_decode_u32(utl_ctx *ctx, const uint8_t *d_src, const uint64_t *d_block_off, const uint64_t *d_out_off,
                   size_t nblocks, uint32_t *d_out, int32_t *d_status, int32_t *d_consumed);

/* Batched UnpackUint64 over an index (one launch). */
int utl_decode_u64(utl_ctx *ctx, const uint8_t *d_src, const uint64_t *d_block_off, const uint64_t *d_out_off,
                   size_t nblocks, uint64_t *d_out, int32_t *d_status, int32_t *d_consumed);

/* Persistent form of the batched decode (is_u64 selects UnpackUint64): a
 * fixed grid of workgroups pulls runs of tiles from a device work queue.
 * Results identical to utl_decode_u32/u64. On gfx950 it is slower than the
 * flat launch (a looping wavefront waits for its previous stores with every
 * load): use it where a fixed grid is wanted. */
int utl_decode_persistent(utl_ctx *ctx, int is_u64, const uint8_t *d_src, const uint64_t *d_block_off,
                          const uint64_t *d_out_off, size_t nblocks, void *d_out, int32_t *d_status,
                          int32_t *d_consumed);

/* Launch variant of the batched decode kernels on a context (no reference
 * counterpart; the reference tunes its own path constants offline,
 * internal/threshold/main.go:36). Chosen per workload by the autotuner
 * (utlpfor_amd/autotune.py); every variant gives identical results.
 *   arena_bytes    LDS bytes a wavefront stages its eight uint32 blocks in:
 *                  2560, 3328 or 4352. Larger arenas take fewer staging
 *                  rounds on streams of wide blocks, at fewer waves per CU.
 *                  uint64 blocks always use an 8320-byte arena.
 *   waves_per_wg   wavefronts per workgroup, 2 or 4 (8 blocks each).
 *   grid_wgs       persistent form only: workgroups in the grid.
 *   tiles_per_grab persistent form only: tiles of 8 * waves_per_wg blocks a
 *                  workgroup takes per work-queue grab.
 * The flat forms (utl_decode_u32/u64, graphs) use arena_bytes and
 * waves_per_wg; the per-call entry points always use the 2560-byte arena. */
typedef struct utl_tune_params {
  int arena_bytes;
  int waves_per_wg;
  int grid_wgs;
  int tiles_per_grab;
} utl_tune_params;
int utl_set_tune(utl_ctx *ctx, const utl_tune_params *t);
int utl_get_tune(utl_ctx *ctx, utl_tune_params *t);

/* Batched GetUint32 / GetUint64 (get.go:74, uint64.go:364): request i reads
 * position d_req_pos[i] of block d_req_block[i]; with d_req_block NULL,
 * d_req_pos[i] is a position in the decoded stream and its block, the last
 * one whose d_out_off entry is <= the position, is found by a search that
 * starts at position / 128 (d_out_off nondecreasing). d_out receives 4- or
 * 8-byte values, d_status the per-request status word. */
int utl_get_batch(utl_ctx *ctx, int is_u64, const uint8_t *d_src, const uint64_t *d_block_off,
                  const uint64_t *d_out_off, size_t nblocks, const uint64_t *d_req_block, const int64_t *d_req_pos,
                  size_t nreq, void *d_out, int32_t *d_status);

/* Position directory for batched Gets by stream position (no reference
 * counterpart; the reference's callers find a block by walking BlockLength):
 * entry c (UTL_POS_DIR_WORDS uint64 words of d_dir) describes the last block
 * whose d_out_off entry is <= 128c (0 when none) and the block after it:
 * their d_block_off / d_out_off entries, for c < ndir; ndir =
 * ceil(d_out_off[nblocks] / 128) entries cover every position of the
 * stream. d_out_off must be nondecreasing; the directory holds both arrays'
 * values and is rebuilt when either changes. */
#define UTL_POS_DIR_WORDS 4
int utl_pos_dir_build(utl_ctx *ctx, const uint64_t *d_block_off, const uint64_t *d_out_off, size_t nblocks,
                      uint64_t *d_dir, size_t ndir);

/* utl_get_batch of stream positions (get.go:74, uint64.go:364) with the
 * block found through d_dir (ndir entries) from utl_pos_dir_build over the
 * same d_block_off and d_out_off: the same values and statuses as
 * utl_get_batch with d_req_block NULL. A position in an entry's block or the
 * next takes their offsets from the entry (one directory read per request);
 * others gallop over d_out_off from the entry's block. */
int utl_get_batch_dir(utl_ctx *ctx, int is_u64, const uint8_t *d_src, const uint64_t *d_block_off,
                      const uint64_t *d_out_off, size_t nblocks, const uint64_t *d_dir, size_t ndir,
                      const int64_t *d_req_pos, size_t nreq, void *d_out, int32_t *d_status);

/* Gathered index for a serving batch (no reference counterpart): request r,
 * row {first block, block count n, value base u, entry base e} of d_req
 * (nreq rows of 4 uint64), fills entries [e, e + n] of d_gblock_off /
 * d_gout_off from the store index d_block_off / d_out_off (store_nblocks
 * blocks), with out offsets rebased to u. Rows must satisfy
 * e_0 = 0, e_{r+1} = e_r + n_r + 1, u_{r+1} = u_r + (values of r), and
 * nentries = e_last + n_last + 1. Decoding the gathered index over
 * nentries - 1 blocks (utl_decode_u32/u64) then decodes every request into
 * out[u_r ..]; the block between two requests has an empty output slot and
 * writes nothing, and its status belongs to no request. */
int utl_gather_index(utl_ctx *ctx, const uint64_t *d_block_off, const uint64_t *d_out_off, size_t store_nblocks,
                     const uint64_t *d_req, size_t nreq, size_t nentries, uint64_t *d_gblock_off,
                     uint64_t *d_gout_off);

/* Batched PackUint32 / PackUint64 (pack.go:76, uint64.go:279) over a job
 * table: block b encodes values[d_in_off[b]..] with
 * d_in_meta[b] = count (1..128) | flag << 8 into d_slots (nblocks *
 * UTL_SLOT_BYTES32/64 bytes): uint64 block b at byte b * UTL_SLOT_BYTES64;
 * uint32 blocks 32t .. 32t + 31 back to back from byte
 * 32t * UTL_SLOT_BYTES32. d_out_meta[b] = length | count << 24, or
 * 0xFFFFFFFF for a count outside 1..128 (ErrInvalidBuffer). utl_compact
 * turns the slots into the stream. */
#define UTL_SLOT_BYTES32 1024u
#define UTL_SLOT_BYTES64 2048u
int utl_encode_blocks(utl_ctx *ctx, int is_u64, const void *d_values, const uint64_t *d_in_off,
                      const uint32_t *d_in_meta, size_t nblocks, uint8_t *d_slots, uint32_t *d_out_meta);
/* Batched PackUint32 straight into the concatenated stream (no slots and no
 * compaction): each workgroup encodes 32 consecutive jobs, learns its byte
 * offset from the ones before it by a decoupled look-back over tiles and
 * writes its blocks at their place in d_dst. Same job table as
 * utl_encode_blocks; d_dst, d_out_meta and d_block_off[0..nblocks] receive
 * what utl_encode_blocks + utl_index_build + utl_compact give, byte for
 * byte (bytes of d_dst past the stream are not written). dst_cap must be at
 * least nblocks * utl_max_block_length32(0). Synchronous: returns the
 * stream's length in *out_nbytes. UTL_ERR_HIP if a tile's look-back did not
 * complete (its waits are bounded; nothing of such a tile is written). */
int utl_encode_u32_direct(utl_ctx *ctx, const uint32_t *d_values, const uint64_t *d_in_off,
                          const uint32_t *d_in_meta, size_t nblocks, uint8_t *d_dst, size_t dst_cap,
                          uint32_t *d_out_meta, uint64_t *d_block_off, size_t *out_nbytes);
/* Batched PackUint32 in two passes with the index scan between them (no
 * slots, no compaction, no waits between workgroups). Same job table as
 * utl_encode_blocks. utl_encode_u32_size makes every block's decisions
 * (FOR, zigzag, bit width, exception count: packBlockScalar, scalar.go:126-235)
 * and writes d_out_meta[b] (length | count << 24, or 0xFFFFFFFF for a count
 * outside 1..128) and d_plan[b] (8 bytes: the decisions). The caller scans
 * d_out_meta (utl_index_build); utl_encode_u32_write then packs block b at
 * d_block_off[b] of d_dst. d_dst, d_out_meta and d_block_off receive what
 * utl_encode_blocks + utl_index_build + utl_compact give, byte for byte. A
 * block that would end past dst_cap is not written. */
int utl_encode_u32_size(utl_ctx *ctx, const uint32_t *d_values, const uint64_t *d_in_off,
                        const uint32_t *d_in_meta, size_t nblocks, uint32_t *d_out_meta, uint64_t *d_plan);
int utl_encode_u32_write(utl_ctx *ctx, const uint32_t *d_values, const uint64_t *d_in_off,
                         const uint32_t *d_in_meta, size_t nblocks, const uint64_t *d_plan,
                         const uint64_t *d_block_off, uint8_t *d_dst, size_t dst_cap);
/* Batched PackUint64 (uint64.go:279) in two passes with the index scan
 * between them: utl_encode_u64_size writes d_out_meta (as
 * utl_encode_blocks does); after utl_index_build over it,
 * utl_encode_u64_write encodes each block again (packUint64Scalar,
 * uint64.go:302-333) and writes it at d_block_off[b] of d_dst: the same
 * bytes as utl_encode_blocks + utl_compact, with no slots. A sub-block that
 * would end past dst_cap is not written. */
int utl_encode_u64_size(utl_ctx *ctx, const uint64_t *d_values, const uint64_t *d_in_off,
                        const uint32_t *d_in_meta, size_t nblocks, uint32_t *d_out_meta);
int utl_encode_u64_write(utl_ctx *ctx, const uint64_t *d_values, const uint64_t *d_in_off,
                         const uint32_t *d_in_meta, size_t nblocks, const uint64_t *d_block_off, uint8_t *d_dst,
                         size_t dst_cap);
/* Concatenates encoded slots into d_dst at d_block_off (from utl_index_build
 * over d_out_meta): byte-identical to appending the blocks with Append.
 * Writes exactly the stream's d_block_off[nblocks] bytes. */
int utl_compact(utl_ctx *ctx, int is_u64, const uint8_t *d_slots, const uint64_t *d_block_off, size_t nblocks,
                uint8_t *d_dst);
/* Job table for streams cut into 128-value blocks (the last one partial):
 * stream s covers values [d_val_off[s], d_val_off[s+1]) with flag d_flags[s],
 * its blocks numbered from d_blk_off[s]. */
int utl_make_jobs(utl_ctx *ctx, const uint64_t *d_val_off, const uint64_t *d_blk_off, const uint8_t *d_flags,
                  size_t nstreams, uint64_t *d_in_off, uint32_t *d_in_meta);

/* One resident stream's whole decode -- the device index scan over d_meta
 * and the batched decode (persistent work-queue form when `persistent`, with
 * the context's current launch variant) -- captured once as a HIP graph and
 * replayed by utl_graph_launch on the context's stream. With d_block_off and
 * d_out_off NULL the graph keeps only the wave index (utl_wave_index_build,
 * in memory it owns) and decodes in meta form (utl_decode_meta); otherwise
 * it writes the full index there and decodes from it. All pointers are
 * fixed at capture. The graph owns the scratch of its index scan, so other
 * calls on the same context never disturb a replay; the persistent form
 * uses the context's work counters, so a graph must not outlive its
 * context. Returns NULL on failure (utl_last_error). */
typedef struct utl_graph utl_graph;
utl_graph *utl_graph_decode_create(utl_ctx *ctx, int is_u64, int persistent, const uint8_t *d_src,
                                   const uint32_t *d_meta, size_t nblocks, uint64_t *d_block_off,
                                   uint64_t *d_out_off, void *d_out, int32_t *d_status);
int utl_graph_launch(utl_graph *g);
void utl_graph_destroy(utl_graph *g);

/* HIP events on the context's stream, for timing device work in place. */
int utl_event_create(void **ev);
int utl_event_destroy(void *ev);
int utl_event_record(utl_ctx *ctx, void *ev);
int utl_event_elapsed_ms(void *start, void *stop, float *ms);
/* Plain device-to-device copy kernel (bandwidth calibration). */
int utl_dev_copy(utl_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
/* A streaming kernel that reads read_bytes from d_src and writes write_bytes
 * to d_dst (16-byte chunks, coalesced; 16-byte aligned pointers) on
 * grid_wgs workgroups of 256 threads (0: four per CU): the practical HBM
 * rate of a kernel with the same traffic (the decode of a stream reads its
 * compressed bytes and writes its values). */
int utl_dev_stream(utl_ctx *ctx, void *d_dst, size_t write_bytes, const void *d_src, size_t read_bytes,
                   int grid_wgs);

/* ---- stream container (no reference counterpart) ----
 *
 * The on-disk / wire form of a set of streams: the blocks exactly as
 * PackUint32/PackUint64(flag|Append, ...) emit them, per 128-value chunk
 * (pack.go:53-76), plus what the index needs and a stream directory:
 *
 *   [utl_container_header, 80 B]
 *   [nstreams x utl_stream_entry, 48 B each]
 *   [nblocks x u32 meta word (byte length | value count << 24), padded to 8 B]
 *   [nbytes of concatenated blocks]
 *
 * All integers are little-endian. Block offsets are not stored: the device
 * scan (utl_index_build) rebuilds them from the meta words. Host-only
 * functions: no device work. */
#define UTL_CONTAINER_MAGIC "UTLPFOR1"
#define UTL_CONTAINER_VERSION 1u

typedef struct utl_container_header {
  char magic[8];         /* UTL_CONTAINER_MAGIC */
  uint32_t version;      /* UTL_CONTAINER_VERSION */
  uint32_t header_bytes; /* 80 */
  uint64_t nstreams, nblocks, nbytes, nvalues;
  uint64_t dir_off, meta_off, data_off, total_bytes;
} utl_container_header;

typedef struct utl_stream_entry {
  uint64_t first_block; /* block range [first_block, first_block + nblocks) */
  uint64_t nblocks;
  uint64_t first_value; /* value range in the decoded concatenation */
  uint64_t nvalues;
  uint64_t byte_off;    /* first byte of the stream in the data section */
  uint32_t flag;        /* the Pack flags the stream was written with */
  uint32_t is_u64;      /* 0: uint32 blocks, 1: uint64 blocks */
} utl_stream_entry;

/* Bytes of a container holding these counts, or 0 on overflow. */
size_t utl_container_bytes(size_t nstreams, size_t nblocks, size_t nbytes);

/* Writes a container into dst (dst_cap >= utl_container_bytes(...)). The
 * directory must tile [0, nblocks) and [0, nbytes) in order, and the meta
 * words of each stream must sum to its bytes and values; UTL_ERR_ARGUMENT
 * otherwise. */
int utl_container_write(uint8_t *dst, size_t dst_cap, const utl_stream_entry *dir, size_t nstreams,
                        const uint32_t *meta, size_t nblocks, const uint8_t *data, size_t nbytes,
                        size_t *out_len);

/* Validates a container held in src[0:len] (header, offsets, directory
 * tiling, meta words against the directory) and returns its header;
 * *dir, *meta and *data point into src. UTL_ERR_INVALID_BUFFER on any
 * inconsistency. The blocks themselves are checked by decoding them. */
int utl_container_parse(const uint8_t *src, size_t len, utl_container_header *hdr, const utl_stream_entry **dir,
                        const uint32_t **meta, const uint8_t **data);

/* Message of the last UTL_ERR_HIP / UTL_ERR_NO_DEVICE on this thread. */
const char *utl_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
