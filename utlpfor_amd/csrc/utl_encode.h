// utl_encode.h -- host-side launchers of the encode kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Job table: block b encodes values[in_off[b] .. +count) with
// in_meta[b] = count (1..128) | flag << 8; out_meta[b] = length | count << 24.
// uint32: the blocks of tile t (blocks 32t .. 32t + 31) are written back to
// back from byte 32t * slot_bytes; uint64: block b at byte b * slot_bytes.
hipError_t utl_launch_encode_u32(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                 uint64_t nblocks, uint8_t* slots, uint32_t slot_bytes, uint32_t* out_meta,
                                 uint32_t* work_out, hipStream_t stream);
// Direct form of the u32 encode: blocks written straight to their place in
// the concatenated stream dst (byte-identical to slots + compaction), with
// their offsets in block_off[0..nblocks]. lb is scratch of
// utl_encode_direct_tmp_bytes(nblocks) bytes; after the launch its last word
// is nonzero if a tile's look-back failed (nothing of that tile written) or
// a block would have ended past dst_cap (that block not written).
size_t utl_encode_direct_tmp_bytes(uint64_t nblocks);
hipError_t utl_launch_encode_u32_direct(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                        uint64_t nblocks, uint8_t* dst, uint64_t dst_cap, uint32_t* out_meta,
                                        uint64_t* block_off, uint32_t* work_out, uint64_t* lb, hipStream_t stream);
// Sized form of the u32 encode: the size pass writes out_meta and plan (8
// bytes a block); after the caller's index scan of out_meta, the write pass
// packs block b at block_off[b] of dst (blocks ending past dst_cap are not
// written). Byte-identical to the slot form + compaction.
hipError_t utl_launch_encode_u32_size(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                      uint64_t nblocks, uint32_t* out_meta, uint64_t* plan, hipStream_t stream);
hipError_t utl_launch_encode_u32_write(const uint32_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                       uint64_t nblocks, const uint64_t* plan, const uint64_t* block_off, uint8_t* dst,
                                       uint64_t dst_cap, hipStream_t stream);
hipError_t utl_launch_encode_u64(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                 uint64_t nblocks, uint8_t* slots, uint32_t slot_bytes, uint32_t* out_meta,
                                 hipStream_t stream);
// Sized form of the u64 encode: the size pass writes out_meta only; after
// the caller's index scan of out_meta, the write pass encodes block b again
// and writes it at block_off[b] of dst (a sub-block ending past dst_cap is
// not written). Byte-identical to the slot form + compaction.
hipError_t utl_launch_encode_u64_size(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                      uint64_t nblocks, uint32_t* out_meta, hipStream_t stream);
hipError_t utl_launch_encode_u64_write(const uint64_t* values, const uint64_t* in_off, const uint32_t* in_meta,
                                       uint64_t nblocks, const uint64_t* block_off, uint8_t* dst, uint64_t dst_cap,
                                       hipStream_t stream);
// Concatenation of encoded slots: tiles (the uint32 layout) or one block per
// slot (uint64), into dst at the block offsets boff.
hipError_t utl_launch_compact(const uint8_t* slots, uint32_t slot_bytes, bool tiles, const uint64_t* boff,
                              uint64_t nblocks, uint8_t* dst, hipStream_t stream);
hipError_t utl_launch_make_jobs(const uint64_t* val_off, const uint64_t* blk_off, const uint8_t* flags,
                                uint64_t nstreams, uint64_t* in_off, uint32_t* in_meta, hipStream_t stream);
