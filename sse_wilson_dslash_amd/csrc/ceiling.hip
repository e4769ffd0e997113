// Streaming ceiling for the Dslash's HBM traffic: a kernel with the access
// mix of the 4D FP32 operator (per target site 6 spinor planes and 36 link
// planes read, the links non-temporally, 6 result planes written, every 16-byte
// plane at the site's own index, so no gathers and no reuse) and the same
// XCD-contiguous block order.  Its time is the practical floor for moving the
// Dslash's compulsory bytes on this GPU; dslash_time_ceiling reports it.
#include <hip/hip_runtime.h>

#include "dslash_common.h"
#include "status.h"

namespace wd {

template <int NS, int NL>
__global__ void __launch_bounds__(256) k_ceiling(v4u* __restrict__ out, const v4u* __restrict__ in,
                                                 const v4u* __restrict__ links, uint32_t vh, int remap) {
  const uint32_t b = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t i = b * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  v4u acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = in[(size_t)k * vh + i];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    acc[k % NS] ^= ld_nt(&links[(size_t)k * vh + i]);
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) out[(size_t)k * vh + i] = acc[k];
}

}  // namespace wd

using namespace wd;

extern "C" {

int dslash_time_ceiling(void* d_out, const void* d_in, const void* d_links, const int latt[4], int block_threads,
                        int xcd_remap, int n, void* stream, float* ms) {
  clear_error();
  WD_REQUIRE(latt && ms && n > 0, "bad arguments");
  if (int e = dslash_check_lattice(latt, 4)) return e;
  WD_REQUIRE(d_out && d_in && d_links, "NULL buffer");
  WD_REQUIRE(block_threads >= 64 && block_threads <= 256 && block_threads % 64 == 0,
             "block_threads must be 64..256, a multiple of 64");
  const uint32_t vh = (uint32_t)((uint64_t)latt[0] * latt[1] * latt[2] * latt[3] / 2);
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((vh + block_threads - 1) / block_threads), blk(block_threads);
  hipEvent_t e0, e1;
  WD_HIP(hipEventCreate(&e0));
  WD_HIP(hipEventCreate(&e1));
  int rc = 0;
  hipError_t he = hipEventRecord(e0, s);
  for (int k = 0; k < n && he == hipSuccess; ++k) {
    hipLaunchKernelGGL((k_ceiling<6, 36>), grid, blk, 0, s, (v4u*)d_out, (const v4u*)d_in, (const v4u*)d_links, vh,
                       xcd_remap);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipEventRecord(e1, s);
  if (he == hipSuccess) he = hipEventSynchronize(e1);
  float t = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
  if (he != hipSuccess) rc = set_error(DSLASH_ERR_HIP, "ceiling kernel: %s", hipGetErrorString(he));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / n;
  return rc;
}

}  // extern "C"
