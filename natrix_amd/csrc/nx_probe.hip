// Copy-bandwidth probe: a plain streaming copy of known size, used to
// measure the attainable HBM bandwidth on the device and to calibrate the
// FETCH_SIZE / WRITE_SIZE hardware counters (one copy of B bytes moves
// exactly B bytes in and B bytes out).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/natrix.h"

namespace {

__global__ void __launch_bounds__(256)
    k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n;
       i += (size_t)gridDim.x * 256)
    dst[i] = src[i];
}

}  // namespace

extern "C" int nx_copy_bandwidth(size_t bytes, int iters, double* gbps) {
  if (!gbps || iters < 1 || bytes < 16 || bytes % 16 != 0) return NX_ERR_ARG;
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = NX_OK;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMemsetAsync(a, 0, bytes, s) != hipSuccess) {
    rc = NX_ERR_HIP;
  } else {
    const unsigned blocks = 256 * 32;  // 32 blocks per CU
    k_copy<<<blocks, 256, 0, s>>>(a, b, n);  // warm-up
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < iters; ++i) k_copy<<<blocks, 256, 0, s>>>(a, b, n);
    (void)hipEventRecord(e1, s);
    float ms = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f)
      rc = NX_ERR_HIP;
    else
      *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
