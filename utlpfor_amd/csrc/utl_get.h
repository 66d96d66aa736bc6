// utl_get.h -- host-side launcher of the batched Get kernel (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Status per request: UTL_* code | overflow position << 8.
hipError_t utl_launch_get(bool u64, const uint8_t* src, const uint64_t* boff, const uint64_t* ooff, uint64_t nblocks,
                          const uint64_t* req_block, const int64_t* req_pos, uint64_t nreq, void* out, int32_t* status,
                          hipStream_t stream);
