// utl_kernels.h -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Launch variant of the decode kernels, chosen by the autotuner.
struct utl_tune {
  int arena_bytes;          // LDS staging bytes per wavefront, uint32 blocks: 2560, 3328 or 4352
  int waves_per_wg;         // wavefronts per workgroup: 2 or 4
  int grid_wgs;             // persistent form: workgroups in the grid
  uint32_t tiles_per_grab;  // persistent form: tiles a workgroup takes from the work queue per grab
};
bool utl_tune_valid(const utl_tune& t);

hipError_t utl_launch_decode(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                             void* out, int32_t* status, int32_t* consumed, const utl_tune& tune, hipStream_t stream);

hipError_t utl_launch_decode_persistent(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff,
                                        uint64_t nblocks, void* out, int32_t* status, int32_t* consumed,
                                        unsigned long long* counter, const utl_tune& tune, hipStream_t stream);

// Decode of concatenated blocks described by their meta words and a wave
// index (the byte and value offsets of every eighth block).
hipError_t utl_launch_decode_meta(bool u64, const uint8_t* src, const uint32_t* meta, const uint64_t* wave,
                                  uint64_t nblocks, void* out, int32_t* status, int32_t* consumed,
                                  unsigned long long* counter, const utl_tune& tune, hipStream_t stream);

// Exclusive scans of (length, count) meta words into block/value offsets.
// tmp must hold utl_index_tmp_bytes(nblocks) bytes.
size_t utl_index_tmp_bytes(uint64_t nblocks);
hipError_t utl_launch_index(const uint32_t* meta, uint64_t nblocks, uint64_t* boff, uint64_t* ooff, void* tmp,
                            hipStream_t stream);

// The index scan's fixed device scratch (utl_index_build_noalloc): its size,
// the once-per-device lookup of its address (with that device current), and
// that address (null before the lookup).
size_t utl_index_static_scratch_bytes();
hipError_t utl_index_scratch_init(int device);
void* utl_index_scratch_ptr(int device);

// The wave index: wave[2w], wave[2w+1] = byte and value offsets of block 8w
// (utl_wave_index_entries(nblocks) entries); same scratch as the index scan.
// meta and wave must be 16-byte aligned.
size_t utl_wave_index_entries(uint64_t nblocks);
hipError_t utl_launch_wave_index(const uint32_t* meta, uint64_t nblocks, uint64_t* wave, void* tmp,
                                 hipStream_t stream);

hipError_t utl_launch_block_lengths(const uint8_t* src, const uint64_t* start, const uint64_t* limit, uint64_t nblocks,
                                    uint32_t* meta, int32_t* status, hipStream_t stream);

// Walk of foreign streams (block starts unknown): counts, then meta words
// (utl_walk.hip). total_bytes = stream_off[nstreams] - stream_off[0] sizes
// the scratch tmp (utl_walk_tmp_bytes bytes). The short streams are walked
// on `side` (forked from and joined back into `stream` with the two
// events), alongside the chunk-parallel walk of the long ones.
size_t utl_walk_tmp_bytes(uint64_t nstreams, uint64_t total_bytes);
hipError_t utl_launch_index_walk(const uint8_t* src, const uint64_t* stream_off, uint64_t nstreams,
                                 uint64_t total_bytes, const uint64_t* blk_off, uint32_t* meta, uint64_t* start,
                                 uint32_t* nblocks, int32_t* status, void* tmp, hipStream_t stream,
                                 hipStream_t side, hipEvent_t ev_fork, hipEvent_t ev_join);

hipError_t utl_launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t stream);
hipError_t utl_launch_fill_bytes(uint8_t* p, uint8_t v, uint64_t n, hipStream_t stream);

// Plain streaming copy (16-byte vectors when both pointers allow it).
hipError_t utl_launch_copy(uint8_t* dst, const uint8_t* src, uint64_t n, hipStream_t stream);
// Streaming read of read_bytes and write of write_bytes (16-byte chunks;
// 16-byte aligned pointers), on `grid` workgroups.
hipError_t utl_launch_stream_rw(uint8_t* dst, uint64_t write_bytes, const uint8_t* src, uint64_t read_bytes,
                                int grid, hipStream_t stream);

// Gathered index of a batch of block ranges of a resident store (utl_serve.hip).
hipError_t utl_launch_gather_index(const uint64_t* boff, const uint64_t* ooff, uint64_t store_nblocks,
                                   const uint64_t* req, uint64_t nreq, uint64_t nentries, uint64_t* gboff,
                                   uint64_t* gooff, hipStream_t stream);
