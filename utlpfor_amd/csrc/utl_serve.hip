// utl_serve.hip -- gathered index for a batch of decompress requests.
//
// A serving batch decodes many block ranges of a resident store in one
// launch. Request r (table row {first block f, block count n, value base u,
// entry base e}) owns entries [e, e + n] of a gathered index:
//   gboff[e + k] = boff[f + k],  gooff[e + k] = u + ooff[f + k] - ooff[f]
// for k = 0..n, so the batched decode over (entries - 1) blocks decodes
// block f + k into out[gooff[e + k] ..). Between two requests the decoder
// sees one extra block, from the end of request r to the start of request
// r + 1; with e_{r+1} = e_r + n_r + 1 and u_{r+1} = u_r + (values of r) its
// output slot holds zero values, so it writes nothing (its status is not a
// request's). This replaces the host-side concatenation of index slices.
#include "utl_device.h"
#include "utl_kernels.h"

namespace utl {

__global__ __launch_bounds__(256) void k_gather_index(const uint64_t* __restrict__ boff,
                                                      const uint64_t* __restrict__ ooff, uint64_t store_nblocks,
                                                      const uint64_t* __restrict__ req, uint64_t nreq,
                                                      uint64_t nentries, uint64_t* __restrict__ gboff,
                                                      uint64_t* __restrict__ gooff) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nentries) return;
  // the last request whose entry base is <= i
  uint64_t lo = 0, hi = nreq - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (req[4 * mid + 3] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint64_t f = req[4 * lo], n = req[4 * lo + 1], u = req[4 * lo + 2], e = req[4 * lo + 3];
  // a table that breaks the layout above still reads inside the store index
  const uint64_t k = min(i - e, n);
  const uint64_t fk = min(f + k, store_nblocks), f0 = min(f, store_nblocks);
  gboff[i] = boff[fk];
  gooff[i] = u + (ooff[fk] - ooff[f0]);
}

}  // namespace utl

hipError_t utl_launch_gather_index(const uint64_t* boff, const uint64_t* ooff, uint64_t store_nblocks,
                                   const uint64_t* req, uint64_t nreq, uint64_t nentries, uint64_t* gboff,
                                   uint64_t* gooff, hipStream_t stream) {
  if (nentries == 0 || nreq == 0) return hipSuccess;
  const unsigned grid = (unsigned)((nentries + 255) / 256);
  utl::k_gather_index<<<grid, 256, 0, stream>>>(boff, ooff, store_nblocks, req, nreq, nentries, gboff, gooff);
  return hipGetLastError();
}
