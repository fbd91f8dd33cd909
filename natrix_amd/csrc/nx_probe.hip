// Copy-bandwidth probe: a plain streaming copy of known size, used to
// measure the attainable HBM bandwidth on the device and to calibrate the
// FETCH_SIZE / WRITE_SIZE hardware counters (one copy of B bytes moves
// exactly B bytes in and B bytes out).
//
// Variants (nx_copy_bandwidth_variant):
//   0  grid-stride, one float4 per thread per iteration (the counter
//      calibration kernel `k_copy`);
//   1  each thread moves U = 4 float4 per iteration, all loads issued
//      before the first store (more bytes in flight per wave);
//   2  U = 8;
//   3  U = 8 with non-temporal loads and stores (the copied data is not
//      kept in the caches);
//   4  one contiguous chunk per workgroup, U = 8 per thread per step
//      (no grid stride: each workgroup streams its own region);
//   5  float2 per lane (8-byte accesses, the stencil kernels' row loads),
//      4 in flight per thread;
//   6  float per lane (4-byte accesses), 4 in flight per thread;
//   7  2-byte accesses (the reduced-storage fields), 4 in flight;
//   8  1-byte accesses (the obstacle masks), 4 in flight.
// The narrower variants calibrate the traffic counters for kernels whose
// loads are narrower than 16 bytes per lane.
// nx_copy_bandwidth runs variant 0; nx_copy_bandwidth_variant(-1) runs all
// of them and reports the best: the attainable rate on this device, against
// which the stencil kernels' measured traffic is read.
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

// U float4 per thread per iteration; element j of a block's iteration is
// base + j * 256 + threadIdx.x, so every load instruction of a wave reads
// one contiguous 1 KiB segment.  n must be a multiple of U * 256 for the
// unguarded body; the tail runs guarded.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ void __launch_bounds__(256)
    k_copy_unrolled(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  const v4f* __restrict__ vs = reinterpret_cast<const v4f*>(src);
  v4f* __restrict__ vd = reinterpret_cast<v4f*>(dst);
  const size_t step = (size_t)gridDim.x * 256 * U;
  for (size_t base = blockIdx.x * (size_t)256 * U; base < n; base += step) {
    v4f v[U];
    if (base + (size_t)256 * U <= n) {
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const v4f* p = vs + base + j * 256 + threadIdx.x;
        v[j] = NT ? __builtin_nontemporal_load(p) : *p;
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        v4f* q = vd + base + j * 256 + threadIdx.x;
        if (NT)
          __builtin_nontemporal_store(v[j], q);
        else
          *q = v[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const size_t i = base + j * 256 + threadIdx.x;
        if (i < n) dst[i] = src[i];
      }
    }
  }
}

// one contiguous chunk of `per_block` float4 per workgroup
template <int U>
__global__ void __launch_bounds__(256)
    k_copy_chunked(const float4* __restrict__ src, float4* __restrict__ dst, size_t n,
                   size_t per_block) {
  const size_t lo = blockIdx.x * per_block;
  const size_t hi = lo + per_block < n ? lo + per_block : n;
  for (size_t base = lo; base < hi; base += (size_t)256 * U) {
    float4 v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const size_t i = base + j * 256 + threadIdx.x;
      v[j] = i < hi ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const size_t i = base + j * 256 + threadIdx.x;
      if (i < hi) dst[i] = v[j];
    }
  }
}

// U elements of type T per thread per iteration, grid-stride (n in T units)
template <class T, int U>
__global__ void __launch_bounds__(256)
    k_copy_narrow(const T* __restrict__ src, T* __restrict__ dst, size_t n) {
  const size_t step = (size_t)gridDim.x * 256 * U;
  for (size_t base = blockIdx.x * (size_t)256 * U; base < n; base += step) {
    T v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const size_t i = base + j * 256 + threadIdx.x;
      if (i < n) v[j] = src[i];
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const size_t i = base + j * 256 + threadIdx.x;
      if (i < n) dst[i] = v[j];
    }
  }
}

constexpr int N_VARIANTS = 9;

void launch_variant(int v, const float4* a, float4* b, size_t n, hipStream_t s) {
  const unsigned blocks = 256 * 32;  // 32 blocks per CU for the grid-stride forms
  switch (v) {
    case 0:
      k_copy<<<blocks, 256, 0, s>>>(a, b, n);
      break;
    case 1:
      k_copy_unrolled<4, false><<<blocks, 256, 0, s>>>(a, b, n);
      break;
    case 2:
      k_copy_unrolled<8, false><<<blocks, 256, 0, s>>>(a, b, n);
      break;
    case 3:
      k_copy_unrolled<8, true><<<blocks, 256, 0, s>>>(a, b, n);
      break;
    case 5:
      k_copy_narrow<float2, 4><<<blocks, 256, 0, s>>>(reinterpret_cast<const float2*>(a),
                                                      reinterpret_cast<float2*>(b), 2 * n);
      break;
    case 6:
      k_copy_narrow<float, 4><<<blocks, 256, 0, s>>>(reinterpret_cast<const float*>(a),
                                                     reinterpret_cast<float*>(b), 4 * n);
      break;
    case 7:
      k_copy_narrow<uint16_t, 4><<<blocks, 256, 0, s>>>(reinterpret_cast<const uint16_t*>(a),
                                                        reinterpret_cast<uint16_t*>(b), 8 * n);
      break;
    case 8:
      k_copy_narrow<uint8_t, 4><<<blocks, 256, 0, s>>>(reinterpret_cast<const uint8_t*>(a),
                                                       reinterpret_cast<uint8_t*>(b), 16 * n);
      break;
    default: {
      // 8 workgroups per CU, each with one contiguous chunk
      const size_t nb = 256 * 8;
      const size_t per = ((n + nb - 1) / nb + 2047) / 2048 * 2048;
      const unsigned g = (unsigned)((n + per - 1) / per);
      k_copy_chunked<8><<<g, 256, 0, s>>>(a, b, n, per);
      break;
    }
  }
}

int copy_bandwidth(size_t bytes, int iters, int variant, double* gbps) {
  if (!gbps || iters < 1 || bytes < 16 || bytes % 16 != 0 || variant < -1 ||
      variant >= N_VARIANTS)
    return NX_ERR_ARG;
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
    double best = 0.0;
    const int lo = variant < 0 ? 0 : variant, hi = variant < 0 ? N_VARIANTS : variant + 1;
    for (int v = lo; v < hi && rc == NX_OK; ++v) {
      launch_variant(v, a, b, n, s);  // warm-up
      (void)hipEventRecord(e0, s);
      for (int i = 0; i < iters; ++i) launch_variant(v, a, b, n, s);
      (void)hipEventRecord(e1, s);
      float ms = 0.0f;
      if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f) {
        rc = NX_ERR_HIP;
      } else {
        const double g = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
        if (g > best) best = g;
      }
    }
    if (rc == NX_OK) *gbps = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

}  // namespace

// the calibration copy (variant 0): FETCH_SIZE / WRITE_SIZE of a run of it
// give the counters' scale
extern "C" int nx_copy_bandwidth(size_t bytes, int iters, double* gbps) {
  return copy_bandwidth(bytes, iters, 0, gbps);
}

extern "C" int nx_copy_bandwidth_variant(size_t bytes, int iters, int variant,
                                         double* gbps) {
  return copy_bandwidth(bytes, iters, variant, gbps);
}
