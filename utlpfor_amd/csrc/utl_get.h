// utl_get.h -- host-side launchers of the batched Get kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Status per request: UTL_* code | overflow position << 8. dir_words (ndir
// entries of 4 words) is the optional position directory of
// utl_launch_pos_dir.
hipError_t utl_launch_get(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                          const uint64_t* dir_words, uint64_t ndir, const uint64_t* req_block, const int64_t* req_pos,
                          uint64_t nreq, void* out, int32_t* status, hipStream_t stream);

// Entry c (4 words) describes the largest b with ooff[b] <= 128c (0 when
// none) and the next block, c < ndir.
hipError_t utl_launch_pos_dir(const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks, uint64_t* dir_words,
                              uint64_t ndir, hipStream_t stream);
