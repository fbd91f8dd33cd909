// FP64 oracle solve of the pressure Poisson system on the GPU.
//
// Conjugate gradients on -A p = -b restricted to fluid cells, where A is the
// operator Jacobi iterates (shader.Poisson.comp:26-37: neighbours outside the
// grid or solid are replaced by the centre).  -A is symmetric positive
// semi-definite on the fluid cells with the constants as null space, so the
// right-hand side is centred on its fluid mean and the residual is
// re-projected every iteration.  All arithmetic and storage are FP64; the
// scalars (rho, alpha, beta) stay in device memory, and the host reads the
// residual every 64 iterations.  Used to measure the error of the float32,
// reduced-storage and learned pressures at full size.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/natrix.h"
#include "nx_kernels.h"

namespace {

constexpr int CG_BLOCKS = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  __syncthreads();
  return t;
}

// q = -A d on fluid cells (0 on solid); part[blk] = sum d.q
__global__ void __launch_bounds__(256)
    k_cg_apply(const double* __restrict__ d, const uint8_t* __restrict__ mask,
               double* __restrict__ q, double* __restrict__ part, int W, int H) {
  __shared__ double sh[4];
  double acc = 0.0;
  const long long n = (long long)W * H;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double v = 0.0;
    if (mask[i] == 0) {
      const int x = (int)(i % W), y = (int)(i / W);
      const double c = d[i];
      const long long iL = (long long)y * W + max(x - 1, 0), iR = (long long)y * W + min(x + 1, W - 1);
      const long long iB = (long long)max(y - 1, 0) * W + x, iT = (long long)min(y + 1, H - 1) * W + x;
      const double x1 = mask[iL] ? c : d[iL];
      const double x2 = mask[iR] ? c : d[iR];
      const double y1 = mask[iB] ? c : d[iB];
      const double y2 = mask[iT] ? c : d[iT];
      v = 4.0 * c - (((x1 + x2) + y1) + y2);
      acc += c * v;
    }
    q[i] = v;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// fixed-order sum of `nb` partials into out[0] (+ optional second array)
__global__ void k_cg_sum(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[4];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) a += part[i];
  const double t = block_sum(a, sh);
  if (threadIdx.x == 0) out[0] = t;
}

// x += alpha d, r -= alpha q (fluid), alpha = s[0] / s[1]; part = sum r, r.r
__global__ void __launch_bounds__(256)
    k_cg_update(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ d,
                const double* __restrict__ q, const uint8_t* __restrict__ mask,
                const double* __restrict__ s, double* __restrict__ part, long long n) {
  __shared__ double sh[4];
  const double alpha = s[1] != 0.0 ? s[0] / s[1] : 0.0;
  double a = 0.0, b = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (mask[i]) continue;
    x[i] += alpha * d[i];
    const double rv = r[i] - alpha * q[i];
    r[i] = rv;
    a += rv;
    b += rv * rv;
  }
  const double ta = block_sum(a, sh);
  const double tb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ta;
    part[CG_BLOCKS + blockIdx.x] = tb;
  }
}

// s layout: [0] rho = r.r, [1] d.q, [2] sum r, [3] r.r (new), [4] fluid count,
// [5] |b|^2 (centred)
__global__ void k_cg_scalars(double* __restrict__ s) {
  // after the update: project r's mean out of rho (r is re-centred next),
  // beta = rho_new / rho
  const double n = s[4];
  const double mean = n > 0 ? s[2] / n : 0.0;
  const double rho_new = s[3] - mean * s[2];
  s[6] = s[0] != 0.0 ? rho_new / s[0] : 0.0;  // beta
  s[7] = mean;
  s[0] = rho_new;
}

// r -= mean (fluid); d = r + beta d (fluid)
__global__ void __launch_bounds__(256)
    k_cg_dir(double* __restrict__ r, double* __restrict__ d, const uint8_t* __restrict__ mask,
             const double* __restrict__ s, long long n) {
  const double beta = s[6], mean = s[7];
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (mask[i]) continue;
    const double rv = r[i] - mean;
    r[i] = rv;
    d[i] = rv + beta * d[i];
  }
}

// init: r = -(b - mean_b) on fluid, x = 0, d = r; part = sum b, sum b^2, count
__global__ void __launch_bounds__(256)
    k_cg_init_sums(const float* __restrict__ b, const uint8_t* __restrict__ mask,
                   double* __restrict__ part, long long n) {
  __shared__ double sh[4];
  double a = 0.0, c = 0.0, cnt = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (mask[i]) continue;
    const double v = (double)b[i];
    a += v;
    c += v * v;
    cnt += 1.0;
  }
  const double ta = block_sum(a, sh), tc = block_sum(c, sh), tn = block_sum(cnt, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ta;
    part[CG_BLOCKS + blockIdx.x] = tc;
    part[2 * CG_BLOCKS + blockIdx.x] = tn;
  }
}

__global__ void __launch_bounds__(256)
    k_cg_init(const float* __restrict__ b, const uint8_t* __restrict__ mask,
              double* __restrict__ x, double* __restrict__ r, double* __restrict__ d,
              double* __restrict__ s, long long n) {
  const double mean = s[8];
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    x[i] = 0.0;
    const double rv = mask[i] ? 0.0 : -((double)b[i] - mean);
    r[i] = rv;
    d[i] = rv;
  }
}

__global__ void k_cg_init_scalars(double* __restrict__ s) {
  const double n = s[4];
  const double mean = n > 0 ? s[2] / n : 0.0;
  s[5] = s[3] - mean * s[2];  // |b - mean|^2
  s[0] = s[5];                // rho = r.r
  s[8] = mean;                // mean of b over fluid
}

// true residual of x: t = (b - mean_b) - A x = (b - mean_b) + q with
// q = -A x; part = sum t, sum t^2 (fluid)
__global__ void __launch_bounds__(256)
    k_cg_true(const float* __restrict__ b, const double* __restrict__ q,
              const uint8_t* __restrict__ mask, const double* __restrict__ s,
              double* __restrict__ part, long long n) {
  __shared__ double sh[4];
  const double mean = s[8];
  double a = 0.0, c = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (mask[i]) continue;
    const double t = ((double)b[i] - mean) + q[i];
    a += t;
    c += t * t;
  }
  const double ta = block_sum(a, sh), tc = block_sum(c, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ta;
    part[CG_BLOCKS + blockIdx.x] = tc;
  }
}

// relative error of a float32 pressure against x, both mean-free on fluid
__global__ void __launch_bounds__(256)
    k_cg_err(const float* __restrict__ p, const double* __restrict__ x,
             const uint8_t* __restrict__ mask, double* __restrict__ part, long long n) {
  __shared__ double sh[4];
  double sp = 0.0, sx = 0.0, spp = 0.0, sxx = 0.0, spx = 0.0, cnt = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (mask[i]) continue;
    const double a = (double)p[i], b = x[i];
    sp += a;
    sx += b;
    spp += a * a;
    sxx += b * b;
    spx += a * b;
    cnt += 1.0;
  }
  double v[6] = {sp, sx, spp, sxx, spx, cnt};
  for (int j = 0; j < 6; ++j) {
    const double t = block_sum(v[j], sh);
    if (threadIdx.x == 0) part[j * CG_BLOCKS + blockIdx.x] = t;
  }
}

}  // namespace

namespace nx {
int set_error(int code, const std::string& msg);
}

#define CG_HIP(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      rc = nx::set_error(NX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
      goto done;                                                           \
    }                                                                      \
  } while (0)

// The solve on caller-provided scratch: r, d, q (n doubles each), part
// (6 CG_BLOCKS doubles) and s (16 doubles).  Every scratch element is
// written by a kernel before it is read, so the scratch may hold anything.
static int fp64_solve_impl(const float* b, const uint8_t* mask, int W, int H, double* x,
                           double tol, int max_iter, double* out, hipStream_t st, double* r,
                           double* d, double* q, double* part, double* s) {
  const long long n = (long long)W * H;
  double hs[16];
  int rc = NX_OK, it = 0;
  const unsigned nb = CG_BLOCKS;
  k_cg_init_sums<<<nb, 256, 0, st>>>(b, mask, part, n);
  k_cg_sum<<<1, 256, 0, st>>>(part, nb, s + 2);
  k_cg_sum<<<1, 256, 0, st>>>(part + CG_BLOCKS, nb, s + 3);
  k_cg_sum<<<1, 256, 0, st>>>(part + 2 * CG_BLOCKS, nb, s + 4);
  k_cg_init_scalars<<<1, 1, 0, st>>>(s);
  k_cg_init<<<nb, 256, 0, st>>>(b, mask, x, r, d, s, n);
  CG_HIP(hipGetLastError());
  CG_HIP(hipMemcpyAsync(hs, s, sizeof(hs), hipMemcpyDeviceToHost, st));
  CG_HIP(hipStreamSynchronize(st));
  {
    const double bnorm2 = hs[5];
    if (bnorm2 <= 0.0) {
      out[0] = 0.0;
      out[1] = 0;
      goto done;
    }
    for (it = 0; it < max_iter; ++it) {
      k_cg_apply<<<nb, 256, 0, st>>>(d, mask, q, part, W, H);
      k_cg_sum<<<1, 256, 0, st>>>(part, nb, s + 1);
      k_cg_update<<<nb, 256, 0, st>>>(x, r, d, q, mask, s, part, n);
      k_cg_sum<<<1, 256, 0, st>>>(part, nb, s + 2);
      k_cg_sum<<<1, 256, 0, st>>>(part + CG_BLOCKS, nb, s + 3);
      k_cg_scalars<<<1, 1, 0, st>>>(s);
      k_cg_dir<<<nb, 256, 0, st>>>(r, d, mask, s, n);
      if ((it & 63) == 63 || it == max_iter - 1) {
        CG_HIP(hipGetLastError());
        CG_HIP(hipMemcpyAsync(hs, s, sizeof(hs), hipMemcpyDeviceToHost, st));
        CG_HIP(hipStreamSynchronize(st));
        if (sqrt(fmax(hs[0], 0.0) / bnorm2) <= tol) {
          ++it;
          break;
        }
      }
    }
    // the true residual of x (not the recursively updated one)
    k_cg_apply<<<nb, 256, 0, st>>>(x, mask, q, part, W, H);
    k_cg_true<<<nb, 256, 0, st>>>(b, q, mask, s, part, n);
    k_cg_sum<<<1, 256, 0, st>>>(part, nb, s + 9);
    k_cg_sum<<<1, 256, 0, st>>>(part + CG_BLOCKS, nb, s + 10);
    CG_HIP(hipGetLastError());
    double h2[16];
    CG_HIP(hipMemcpyAsync(h2, s, sizeof(h2), hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
    const double num = fmax(h2[10] - h2[9] * h2[9] / fmax(h2[4], 1.0), 0.0);
    out[0] = sqrt(num / bnorm2);
    out[1] = it;
  }
done:
  return rc;
}

static bool fp64_solve_args_ok(const float* b, const uint8_t* mask, int W, int H,
                               const double* x, double tol, int max_iter, const double* out) {
  return b && mask && x && out && W >= 1 && H >= 1 && tol > 0 && max_iter >= 1;
}

// Scratch of the allocation-free solve (nx_fp64_solve_noalloc): r, d, q for
// up to NOALLOC_MAX_CELLS cells, then part and s; about 6 MiB, held by every
// device the library runs on until the process ends.  Host code never uses
// this name as an address: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device (noalloc_scratch).
constexpr long long NOALLOC_MAX_CELLS = 1ll << 18;  // 512 x 512
constexpr long long NOALLOC_DOUBLES = 3 * NOALLOC_MAX_CELLS + 6 * CG_BLOCKS + 16;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[NOALLOC_MAX_DEVICES] = {};

// the current device's address of `scratch`, looked up on the first call
// on that device
static int noalloc_scratch(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return nx::set_error(NX_ERR_HIP, "no usable current device for the allocation-free solve");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nx::set_error(NX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return NX_OK;
}

extern "C" {

// FP64 CG solve of A x = b (b: float32 device field, mean removed) on the
// fluid cells of `mask`; x (device, W*H doubles, 0 on solid) gets the
// mean-free solution.  out[0] = relative residual |P(b - A x)| / |P b|,
// out[1] = iterations.
int nx_fp64_solve(const float* b, const uint8_t* mask, int W, int H, double* x,
                  double tol, int max_iter, double* out, void* stream) {
  if (!fp64_solve_args_ok(b, mask, W, H, x, tol, max_iter, out))
    return nx::set_error(NX_ERR_ARG, "bad fp64 solve arguments");
  const long long n = (long long)W * H;
  double *r = nullptr, *d = nullptr, *q = nullptr, *part = nullptr, *s = nullptr;
  int rc = NX_OK;
  CG_HIP(hipMalloc(&r, n * sizeof(double)));
  CG_HIP(hipMalloc(&d, n * sizeof(double)));
  CG_HIP(hipMalloc(&q, n * sizeof(double)));
  CG_HIP(hipMalloc(&part, 6 * CG_BLOCKS * sizeof(double)));
  CG_HIP(hipMalloc(&s, 16 * sizeof(double)));
  rc = fp64_solve_impl(b, mask, W, H, x, tol, max_iter, out, (hipStream_t)stream, r, d, q,
                       part, s);
done:
  (void)hipFree(r);
  (void)hipFree(d);
  (void)hipFree(q);
  (void)hipFree(part);
  (void)hipFree(s);
  return rc;
}

// nx_fp64_solve on the library's static device scratch: no allocation.
int nx_fp64_solve_noalloc(const float* b, const uint8_t* mask, int W, int H, double* x,
                          double tol, int max_iter, double* out, void* stream) {
  if (!fp64_solve_args_ok(b, mask, W, H, x, tol, max_iter, out))
    return nx::set_error(NX_ERR_ARG, "bad fp64 solve arguments");
  const long long n = (long long)W * H;
  if (n > NOALLOC_MAX_CELLS)
    return nx::set_error(NX_ERR_ARG,
                         "nx_fp64_solve_noalloc: " + std::to_string(W) + " x " +
                             std::to_string(H) + " exceeds its scratch of " +
                             std::to_string(NOALLOC_MAX_CELLS) + " cells (512 x 512)");
  double* base = nullptr;
  const int rc = noalloc_scratch(&base);
  if (rc != NX_OK) return rc;
  double* r = base;
  double* d = r + NOALLOC_MAX_CELLS;
  double* q = d + NOALLOC_MAX_CELLS;
  double* part = q + NOALLOC_MAX_CELLS;
  double* s = part + 6 * CG_BLOCKS;
  return fp64_solve_impl(b, mask, W, H, x, tol, max_iter, out, (hipStream_t)stream, r, d, q,
                         part, s);
}

// relative L2 error of a float32 pressure p against the FP64 solution x,
// both with their fluid means removed: |P p - P x| / |P x|
int nx_fp64_compare(const float* p, const double* x, const uint8_t* mask, int W, int H,
                    double* rel, void* stream) {
  if (!p || !x || !mask || !rel) return nx::set_error(NX_ERR_ARG, "null argument");
  const long long n = (long long)W * H;
  const hipStream_t st = (hipStream_t)stream;
  double* part = nullptr;
  double* s = nullptr;
  double h[6];
  int rc = NX_OK;
  CG_HIP(hipMalloc(&part, 6 * CG_BLOCKS * sizeof(double)));
  CG_HIP(hipMalloc(&s, 6 * sizeof(double)));
  k_cg_err<<<CG_BLOCKS, 256, 0, st>>>(p, x, mask, part, n);
  for (int j = 0; j < 6; ++j) k_cg_sum<<<1, 256, 0, st>>>(part + j * CG_BLOCKS, CG_BLOCKS, s + j);
  CG_HIP(hipGetLastError());
  CG_HIP(hipMemcpyAsync(h, s, sizeof(h), hipMemcpyDeviceToHost, st));
  CG_HIP(hipStreamSynchronize(st));
  {
    const double cnt = fmax(h[5], 1.0);
    // |Pp - Px|^2 = (spp - sp^2/n) + (sxx - sx^2/n) - 2 (spx - sp sx / n)
    const double pp = h[2] - h[0] * h[0] / cnt, xx = h[3] - h[1] * h[1] / cnt;
    const double px = h[4] - h[0] * h[1] / cnt;
    *rel = xx > 0 ? sqrt(fmax(pp + xx - 2.0 * px, 0.0) / xx) : 0.0;
  }
done:
  (void)hipFree(part);
  (void)hipFree(s);
  return rc;
}

}  // extern "C"
