// Per-stage fluid kernels for gfx950: one kernel per reference shader.
//
// These are the exact transcriptions used by the reference-order step and by
// the per-stage entry points; the fused fast path lives in nx_fused.hip and
// nx_jacobi.hip.  All are memory-bound pointwise or 5-point-stencil kernels:
// a 64x4 block puts one wavefront on 64 consecutive x of one row, so every
// float2 access of a wavefront is one contiguous 512-byte segment.
#include "nx_kernels.h"

namespace nx {

namespace {

constexpr int BX = 64, BY = 4;

// blocks of BX x BY cells in XCD order (cell_grid / cell_of)
inline dim3 grid2d(int W, int H) { return cell_grid(W, H); }

__device__ __forceinline__ bool cell(int& x, int& y, int W, int H) {
  return cell_of(x, y, W, H);
}

// ---------------------------------------------------------------- S1
// shader.InitBoundaries.comp:18-34 (+ the build's free-slip / inflow modes)
__global__ void __launch_bounds__(256) k_bc(float2* u, int W, int H,
                                            BcParams bc) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const bool ring = x == 0 || y == 0 || x == W - 1 || y == H - 1;
  if (!ring) return;
  // ring values depend only on the cell and interior cells, none of which
  // this kernel writes
  u[y * W + x] = load_bc(u, x, y, W, H, bc);
}

// ---------------------------------------------------------------- S2
__global__ void __launch_bounds__(256)
    k_advect(const float2* __restrict__ uin, float2* __restrict__ uout,
             const uint8_t* __restrict__ mask, int W, int H, float dt,
             float speed, float diss, BcParams bc) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  uout[y * W + x] = advect_cell(uin, mask, x, y, W, H, dt, speed, diss, bc);
}

// ---------------------------------------------------------------- S3
// shader.CalcVorticity.comp:20-26
__global__ void __launch_bounds__(256)
    k_curl(const float2* __restrict__ u, float* __restrict__ w, int W, int H) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const float2 vL = u[y * W + max(x - 1, 0)];
  const float2 vR = u[y * W + min(x + 1, W - 1)];
  const float2 vB = u[max(y - 1, 0) * W + x];
  const float2 vT = u[min(y + 1, H - 1) * W + x];
  w[y * W + x] = 0.5f * ((vR.y - vL.y) - (vT.x - vB.x));
}

// ---------------------------------------------------------------- S4
// shader.ApplyVorticity.comp:26-39
__device__ __forceinline__ float2 confine(float2 u, float wc, float wl,
                                          float wr, float wb, float wt,
                                          float dt, float scale) {
  float fx = 0.5f * (fabsf(wt) - fabsf(wb));
  float fy = 0.5f * (fabsf(wr) - fabsf(wl));
  const float mag = fmaxf(VORTICITY_EPSILON, fx * fx + fy * fy);
  const float inv = 1.0f / sqrtf(mag);  // inversesqrt
  fx = fx * inv;
  fy = fy * inv;
  const float s = scale * wc;  // _VorticityScale * vC * vec2(1, -1)
  fx = fx * s;
  fy = fy * (-s);
  return make_float2(u.x + fx * dt, u.y + fy * dt);
}

__global__ void __launch_bounds__(256)
    k_confine(const float2* __restrict__ uin, const float* __restrict__ w,
              float2* __restrict__ uout, int W, int H, float dt, float scale) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int pos = y * W + x;
  uout[pos] = confine(uin[pos], w[pos], w[y * W + max(x - 1, 0)],
                      w[y * W + min(x + 1, W - 1)], w[max(y - 1, 0) * W + x],
                      w[min(y + 1, H - 1) * W + x], dt, scale);
}

// ---------------------------------------------------------------- S5
// shader.Viscosity.comp:24-31
__global__ void __launch_bounds__(256)
    k_viscosity(const float2* __restrict__ uin, float2* __restrict__ uout,
                int W, int H, float alpha, float rbeta) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const float2 L = uin[y * W + max(x - 1, 0)];
  const float2 R = uin[y * W + min(x + 1, W - 1)];
  const float2 B = uin[max(y - 1, 0) * W + x];
  const float2 T = uin[min(y + 1, H - 1) * W + x];
  const float2 c = uin[y * W + x];
  uout[y * W + x] =
      make_float2(((((L.x + R.x) + B.x) + T.x) + c.x * alpha) * rbeta,
                  ((((L.y + R.y) + B.y) + T.y) + c.y * alpha) * rbeta);
}

// ---------------------------------------------------------------- S6
// shader.Divergence.comp:22-40
__global__ void __launch_bounds__(256)
    k_divergence(const float2* __restrict__ u,
                 const uint8_t* __restrict__ mask, float* __restrict__ div,
                 int W, int H) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int iL = y * W + max(x - 1, 0), iR = y * W + min(x + 1, W - 1);
  const int iB = max(y - 1, 0) * W + x, iT = min(y + 1, H - 1) * W + x;
  const float x1 = mask[iL] ? 0.0f : u[iL].x;
  const float x2 = mask[iR] ? 0.0f : u[iR].x;
  const float y1 = mask[iB] ? 0.0f : u[iB].y;
  const float y2 = mask[iT] ? 0.0f : u[iT].y;
  div[y * W + x] = 0.5f * ((x2 - x1) + (y2 - y1));
}

// ---------------------------------------------------------------- S8
// shader.Poisson.comp:24-37: one sweep
__global__ void __launch_bounds__(256)
    k_jacobi(const float* __restrict__ pin, const float* __restrict__ b,
             const uint8_t* __restrict__ mask, float* __restrict__ pout,
             int W, int H) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int pos = y * W + x;
  const int iL = y * W + max(x - 1, 0), iR = y * W + min(x + 1, W - 1);
  const int iB = max(y - 1, 0) * W + x, iT = min(y + 1, H - 1) * W + x;
  const float p = pin[pos];
  const float x1 = mask[iL] ? p : pin[iL];
  const float x2 = mask[iR] ? p : pin[iR];
  const float y1 = mask[iB] ? p : pin[iB];
  const float y2 = mask[iT] ? p : pin[iT];
  pout[pos] = ((((x1 + x2) + y1) + y2) - b[pos]) * 0.25f;
}

// ---------------------------------------------------------------- S9 + S10
// shader.SubtractGradient.comp:24-46; the end-of-step obstacle clear
// (fluid_simulator.py:268-280) is written to the other mask buffer so the
// stencil reads above never race with it.  The persistent bit survives.
__global__ void __launch_bounds__(256)
    k_subgrad(const float2* __restrict__ uin, const float* __restrict__ p,
              const uint8_t* __restrict__ mask, float2* __restrict__ uout,
              uint8_t* __restrict__ mask_next, int W, int H) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int pos = y * W + x;
  uout[pos] = subgrad_cell(uin, p, mask, x, y, W, H);
  if (mask_next) mask_next[pos] = mask[pos] & MASK_PERSIST;
}

__global__ void k_fill(float* __restrict__ a, float v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_mask_keep(uint8_t* __restrict__ m, uint8_t keep, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    m[i] &= keep;
}

// ---------------------------------------------------------------- F1
// shader.AddVelocity.comp:26-35
__global__ void __launch_bounds__(256)
    k_add_velocity(const float2* __restrict__ uin, float2* __restrict__ uout,
                   int W, int H, float px, float py, float vx, float vy,
                   float r) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int pos = y * W + x;
  const float sx = px * (float)W, sy = py * (float)H;
  const float dx = sx - (float)x, dy = sy - (float)y;
  const float len = sqrtf(dx * dx + dy * dy);
  const float2 val = uin[pos];
  float2 res = val;
  if (len <= r) {
    res.x = val.x + (vx * (r - len)) / r;
    res.y = val.y + (vy * (r - len)) / r;
  }
  uout[pos] = make_float2(clampf(res.x, -1.0f, 1.0f), clampf(res.y, -1.0f, 1.0f));
}

// ---------------------------------------------------------------- F2
// shader.AddCircleObstacle.comp:24-36 (both _Static branches write (1,0))
__global__ void __launch_bounds__(256)
    k_circle(uint8_t* __restrict__ mask, int W, int H, float px, float py,
             float r) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const float sx = px * (float)W, sy = py * (float)H;
  const float dx = sx - (float)x, dy = sy - (float)y;
  if (sqrtf(dx * dx + dy * dy) <= r) {
    const int pos = y * W + x;
    mask[pos] = (mask[pos] & MASK_PERSIST) | MASK_DYN_X;
  }
}

// ---------------------------------------------------------------- F3
// shader.AddTriangleObstacle.comp:19-51
__device__ __forceinline__ float tri_sign(float p1x, float p1y, float p2x,
                                          float p2y, float p3x, float p3y) {
  return ((p1x - p3x) * (p2y - p3y)) - ((p2x - p3x) * (p1y - p3y));
}

__global__ void __launch_bounds__(256)
    k_triangle(uint8_t* __restrict__ mask, int W, int H, float x1, float y1,
               float x2, float y2, float x3, float y3, int is_static) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const float ptx = (float)x / (float)W, pty = (float)y / (float)H;
  const bool b1 = tri_sign(ptx, pty, x1, y1, x2, y2) < 0.0f;
  const bool b2 = tri_sign(ptx, pty, x2, y2, x3, y3) < 0.0f;
  const bool b3 = tri_sign(ptx, pty, x3, y3, x1, y1) < 0.0f;
  if ((b1 == b2) && (b2 == b3)) {
    const int pos = y * W + x;
    mask[pos] = (mask[pos] & MASK_PERSIST) | (is_static ? MASK_DYN_Y : MASK_DYN_X);
  }
}

// ---------------------------------------------------------------- S11
// demo/shaders/shader.AdvectParticle.comp:21-69 (dye grid P, velocity grid V)
__global__ void __launch_bounds__(256)
    k_advect_dye(const float* __restrict__ qin, float* __restrict__ qout,
                 const float2* __restrict__ u,
                 const uint8_t* __restrict__ mask, int Wp, int Hp, int Wv,
                 int Hv, float dt, float speed, float diss) {
  int x, y;
  if (!cell(x, y, Wp, Hp)) return;
  qout[y * Wp + x] =
      advect_dye_cell(qin, u, mask, x, y, Wp, Hp, Wv, Hv, dt, speed, diss);
}

// ---------------------------------------------------------------- F4
// demo/shaders/shader.AddParticle.comp:24-34
__global__ void __launch_bounds__(256)
    k_add_dye(const float* __restrict__ qin, float* __restrict__ qout, int W,
              int H, float px, float py, float val, float r) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int pos = y * W + x;
  const float sx = px * (float)W, sy = py * (float)H;
  const float dx = sx - (float)x, dy = sy - (float)y;
  const float len = sqrtf(dx * dx + dy * dy);
  const float q = qin[pos];
  float res = q;
  if (len <= r) res = clampf(q + (val * (r - len)) / r, 0.0f, 255.0f);
  qout[pos] = res;
}

// ---------------------------------------------------------------- e2
// Gaussian splats (the build's own spec, oracle/fluid.py:gaussian_weight):
// w = float(exp(double(-(dx*dx + dy*dy) / (r*r)))), applied in order.
struct SplatBatch {
  int n;
  Splat s[16];
};

__global__ void __launch_bounds__(256)
    k_gaussian(float2* __restrict__ u, float* __restrict__ q, int W, int H,
               SplatBatch sb) {
  int x, y;
  if (!cell(x, y, W, H)) return;
  const int pos = y * W + x;
  float2 v = u[pos];
  float d = q ? q[pos] : 0.0f;
  for (int i = 0; i < sb.n; ++i) {
    const Splat s = sb.s[i];
    const float w = gaussian_weight(s, x, y, W, H);
    v.x = v.x + s.ax * w;
    v.y = v.y + s.ay * w;
    d = d + s.dye * w;
  }
  u[pos] = v;
  if (q) q[pos] = d;
}

}  // namespace

#define NX_LAUNCH(...)          \
  do {                          \
    __VA_ARGS__;                \
    return hipGetLastError();   \
  } while (0)

hipError_t launch_bc(float2* u, int W, int H, BcParams bc, hipStream_t s) {
  NX_LAUNCH(k_bc<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(u, W, H, bc));
}
hipError_t launch_advect(const float2* uin, float2* uout, const uint8_t* mask,
                         int W, int H, float dt, float speed, float diss,
                         BcParams bc, hipStream_t s) {
  NX_LAUNCH(k_advect<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      uin, uout, mask, W, H, dt, speed, diss, bc));
}
hipError_t launch_curl(const float2* u, float* w, int W, int H,
                       hipStream_t s) {
  NX_LAUNCH(k_curl<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(u, w, W, H));
}
hipError_t launch_confine(const float2* uin, const float* w, float2* uout,
                          int W, int H, float dt, float scale, hipStream_t s) {
  NX_LAUNCH(k_confine<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(uin, w, uout, W,
                                                            H, dt, scale));
}
hipError_t launch_viscosity(const float2* uin, float2* uout, int W, int H,
                            float alpha, float rbeta, hipStream_t s) {
  NX_LAUNCH(k_viscosity<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      uin, uout, W, H, alpha, rbeta));
}
hipError_t launch_divergence(const float2* u, const uint8_t* mask, float* div,
                             int W, int H, hipStream_t s) {
  NX_LAUNCH(k_divergence<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(u, mask, div,
                                                               W, H));
}
hipError_t launch_jacobi_sweep(const float* pin, const float* b,
                               const uint8_t* mask, float* pout, int W, int H,
                               hipStream_t s) {
  NX_LAUNCH(k_jacobi<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(pin, b, mask, pout,
                                                           W, H));
}
hipError_t launch_subtract_gradient(const float2* uin, const float* p,
                                    const uint8_t* mask, float2* uout,
                                    uint8_t* mask_next, int W, int H,
                                    hipStream_t s) {
  NX_LAUNCH(k_subgrad<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      uin, p, mask, uout, mask_next, W, H));
}
hipError_t launch_fill(float* a, float v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  NX_LAUNCH(k_fill<<<(unsigned)(blocks < 8192 ? blocks : 8192), 256, 0, s>>>(
      a, v, n));
}
hipError_t launch_mask_keep(uint8_t* m, uint8_t keep, size_t n,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  NX_LAUNCH(k_mask_keep<<<(unsigned)(blocks < 8192 ? blocks : 8192), 256, 0,
                          s>>>(m, keep, n));
}
hipError_t launch_add_velocity(const float2* uin, float2* uout, int W, int H,
                               float px, float py, float vx, float vy,
                               float r, hipStream_t s) {
  NX_LAUNCH(k_add_velocity<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      uin, uout, W, H, px, py, vx, vy, r));
}
hipError_t launch_circle(uint8_t* mask, int W, int H, float px, float py,
                         float r, hipStream_t s) {
  NX_LAUNCH(k_circle<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(mask, W, H, px, py,
                                                           r));
}
hipError_t launch_triangle(uint8_t* mask, int W, int H, float x1, float y1,
                           float x2, float y2, float x3, float y3,
                           int is_static, hipStream_t s) {
  NX_LAUNCH(k_triangle<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      mask, W, H, x1, y1, x2, y2, x3, y3, is_static));
}
hipError_t launch_advect_dye(const float* qin, float* qout, const float2* u,
                             const uint8_t* mask, int Wp, int Hp, int Wv,
                             int Hv, float dt, float speed, float diss,
                             hipStream_t s) {
  NX_LAUNCH(k_advect_dye<<<grid2d(Wp, Hp), dim3(BX, BY), 0, s>>>(
      qin, qout, u, mask, Wp, Hp, Wv, Hv, dt, speed, diss));
}
hipError_t launch_add_dye(const float* qin, float* qout, int W, int H,
                          float px, float py, float val, float r,
                          hipStream_t s) {
  NX_LAUNCH(k_add_dye<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(qin, qout, W, H,
                                                            px, py, val, r));
}
hipError_t launch_gaussian_splats(float2* u, float* q, int W, int H,
                                  const Splat* splats, int n, hipStream_t s) {
  // splats travel by value in the kernel arguments, 16 per launch, so the
  // caller's host array may be reused as soon as this returns
  for (int i0 = 0; i0 < n; i0 += 16) {
    SplatBatch sb;
    sb.n = n - i0 < 16 ? n - i0 : 16;
    for (int i = 0; i < sb.n; ++i) sb.s[i] = splats[i0 + i];
    k_gaussian<<<grid2d(W, H), dim3(BX, BY), 0, s>>>(u, q, W, H, sb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace nx
