// Batched FP64 conjugate gradients for many small pressure problems: the
// training targets of the learned pressure guess.
//
// Each problem is an n x n grid (n * n <= 16384 cells) with its own
// right-hand side b (float32) and solid mask, stored back to back.  One
// workgroup of 1024 threads solves one problem to the end in a single
// launch: thread t owns cells t, t + 1024, ... (at most 16) and keeps x, r
// and d of them in registers; the search direction goes through the
// problem's slice of `work` (device, FP64) so that neighbours can read it,
// with a workgroup barrier between the store and the stencil (the waves of
// a workgroup share their compute unit's caches, so the barrier orders the
// global stores and loads).  The operator is the one Jacobi iterates
// (shader.Poisson.comp:26-37: a solid or out-of-grid neighbour reads the
// centre), restricted to fluid cells; -A is symmetric positive semi-definite
// there with the constants as null space, so b is centred on its fluid mean
// and r is re-centred every iteration (as nx_fp64_solve does on one grid).
// Dot products are workgroup reductions in a fixed order: the result does
// not depend on timing.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "../../include/natrix.h"
#include "nx_kernels.h"

namespace {

constexpr int BCG_THREADS = 1024;
constexpr int BCG_CELLS = 16;  // cells per thread: n * n <= 16384

// sum over the workgroup of two values; every thread gets the totals
__device__ __forceinline__ void block_sum2(double& a, double& b, double (*sh)[2]) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[wave][0] = a;
    sh[wave][1] = b;
  }
  __syncthreads();
  double ta = 0.0, tb = 0.0;
  for (int w = 0; w < BCG_THREADS / 64; ++w) {
    ta += sh[w][0];
    tb += sh[w][1];
  }
  __syncthreads();
  a = ta;
  b = tb;
}

// q = -A v at cell i of an n x n problem (v through global memory), fluid i
__device__ __forceinline__ double neg_lap(const double* __restrict__ v,
                                         const uint8_t* __restrict__ m, int n, int i,
                                         double c) {
  const int x = i % n, y = i / n;
  const int iL = x > 0 ? i - 1 : i, iR = x < n - 1 ? i + 1 : i;
  const int iB = y > 0 ? i - n : i, iT = y < n - 1 ? i + n : i;
  const double x1 = m[iL] ? c : v[iL];
  const double x2 = m[iR] ? c : v[iR];
  const double y1 = m[iB] ? c : v[iB];
  const double y2 = m[iT] ? c : v[iT];
  return 4.0 * c - (((x1 + x2) + y1) + y2);
}

__global__ void __launch_bounds__(BCG_THREADS)
    k_cg_batch(const float* __restrict__ b_all, const uint8_t* __restrict__ m_all,
               double* __restrict__ work, float* __restrict__ xf_all, int n, double tol,
               int max_iter, double* __restrict__ res_out, int* __restrict__ it_out) {
  __shared__ double sh[BCG_THREADS / 64][2];
  const int nn = n * n;
  const long long base = (long long)blockIdx.x * nn;
  const float* b = b_all + base;
  const uint8_t* m = m_all + base;
  double* D = work + base;
  float* xf = xf_all + base;
  const int t = threadIdx.x;

  // per cell: x, r, d in registers (q = -A d is evaluated twice per
  // iteration rather than kept, and b is read again at the end)
  double x[BCG_CELLS], r[BCG_CELLS], d[BCG_CELLS];
  bool fl[BCG_CELLS];
  double sb = 0.0, cnt = 0.0;
#pragma unroll
  for (int j = 0; j < BCG_CELLS; ++j) {
    const int i = t + BCG_THREADS * j;
    fl[j] = i < nn && m[i] == 0;
    sb += fl[j] ? (double)b[i] : 0.0;
    cnt += fl[j] ? 1.0 : 0.0;
  }
  block_sum2(sb, cnt, sh);
  const double mean_b = cnt > 0.0 ? sb / cnt : 0.0;
  double bb = 0.0, unused = 0.0;
#pragma unroll
  for (int j = 0; j < BCG_CELLS; ++j) {
    const int i = t + BCG_THREADS * j;
    const double bcj = fl[j] ? (double)b[i] - mean_b : 0.0;
    bb += bcj * bcj;
    x[j] = 0.0;
    r[j] = -bcj;  // CG on -A x = -b
    d[j] = r[j];
  }
  block_sum2(bb, unused, sh);
  double rho = bb;
  int it = 0;
  if (bb > 0.0) {
    for (; it < max_iter; ++it) {
#pragma unroll
      for (int j = 0; j < BCG_CELLS; ++j) {
        const int i = t + BCG_THREADS * j;
        if (i < nn) D[i] = d[j];
      }
      __syncthreads();
      double dq = 0.0, z = 0.0;
#pragma unroll
      for (int j = 0; j < BCG_CELLS; ++j) {
        const int i = t + BCG_THREADS * j;
        if (fl[j]) dq += d[j] * neg_lap(D, m, n, i, d[j]);
      }
      block_sum2(dq, z, sh);
      const double alpha = dq != 0.0 ? rho / dq : 0.0;
      double sr = 0.0, srr = 0.0;
#pragma unroll
      for (int j = 0; j < BCG_CELLS; ++j) {
        if (!fl[j]) continue;
        const int i = t + BCG_THREADS * j;
        x[j] += alpha * d[j];
        r[j] -= alpha * neg_lap(D, m, n, i, d[j]);
        sr += r[j];
        srr += r[j] * r[j];
      }
      block_sum2(sr, srr, sh);
      const double mean_r = cnt > 0.0 ? sr / cnt : 0.0;
      const double rho_new = srr - mean_r * sr;
      const double beta = rho != 0.0 ? rho_new / rho : 0.0;
#pragma unroll
      for (int j = 0; j < BCG_CELLS; ++j) {
        if (!fl[j]) continue;
        r[j] -= mean_r;
        d[j] = r[j] + beta * d[j];
      }
      rho = rho_new;
      if (sqrt(fmax(rho, 0.0) / bb) <= tol) {
        ++it;
        break;
      }
    }
  }
  // the true residual of x: t = b_c - A x = b_c + (-A x), mean-free
  __syncthreads();
#pragma unroll
  for (int j = 0; j < BCG_CELLS; ++j) {
    const int i = t + BCG_THREADS * j;
    if (i < nn) D[i] = x[j];
  }
  __syncthreads();
  double st = 0.0, stt = 0.0, sx = 0.0, z2 = 0.0;
#pragma unroll
  for (int j = 0; j < BCG_CELLS; ++j) {
    const int i = t + BCG_THREADS * j;
    if (!fl[j]) continue;
    const double tv = ((double)b[i] - mean_b) + neg_lap(D, m, n, i, x[j]);
    st += tv;
    stt += tv * tv;
    sx += x[j];
  }
  block_sum2(st, stt, sh);
  block_sum2(sx, z2, sh);
  const double mean_x = cnt > 0.0 ? sx / cnt : 0.0;
#pragma unroll
  for (int j = 0; j < BCG_CELLS; ++j) {
    const int i = t + BCG_THREADS * j;
    if (i < nn) xf[i] = fl[j] ? (float)(x[j] - mean_x) : 0.0f;
  }
  if (t == 0) {
    const double num = fmax(stt - (cnt > 0.0 ? st * st / cnt : 0.0), 0.0);
    res_out[blockIdx.x] = bb > 0.0 ? sqrt(num / bb) : 0.0;
    it_out[blockIdx.x] = it;
  }
}

}  // namespace

extern "C" int nx_fp64_solve_batch(const float* b, const uint8_t* mask, int n, int count,
                                   double* work, float* xf, double tol, int max_iter,
                                   double* residuals, int* iterations, void* stream) {
  if (!b || !mask || !work || !xf || !residuals || !iterations || n < 1 ||
      n * n > BCG_THREADS * BCG_CELLS || count < 1 || !(tol > 0) || max_iter < 1)
    return nx::set_error(NX_ERR_ARG,
                         "fp64 batch solve: n * n <= 16384, count >= 1, tol > 0, max_iter >= 1");
  k_cg_batch<<<count, BCG_THREADS, 0, (hipStream_t)stream>>>(b, mask, work, xf, n, tol,
                                                             max_iter, residuals, iterations);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return nx::set_error(NX_ERR_HIP, std::string("fp64 batch solve: ") + hipGetErrorString(e));
  return NX_OK;
}
