// Reduced-precision storage: the fused step with velocity, pressure,
// divergence and dye stored as bfloat16 or float16 (half the bytes of
// float32) and every stencil evaluated in float32 registers.
//
// The kernels are the float32 path's device code instantiated with a
// storage policy (nx_common.h): loads widen, stores round to nearest even.
// After the Jacobi solve, a residual-correction launch computes the float32
// residual of the stored (rounded) iterate, runs Jacobi on the correction
// equation from zero in registers, and writes the float32 pressure p + e
// that the projection consumes.
#include "nx_fused_tile.h"
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"

namespace nx {

namespace {

constexpr int BX = 64, BY = 4;

// blocks of BX x BY cells in XCD order (cell_grid / cell_of)
inline dim3 grid2d(int W, int H) { return cell_grid(W, H); }

template <class S>
__global__ void __launch_bounds__(256)
    km_advect(const typename S::s2* __restrict__ uin, typename S::s2* __restrict__ uout,
              const uint8_t* __restrict__ mask, int W, int H, float dt,
              float speed, float diss, BcParams bc) {
  int x, y;
  cell_of(x, y, W, H);
  if (x >= W || y >= H) return;
  S::st2(uout, y * W + x,
         advect_cell<S>(uin, mask, x, y, W, H, dt, speed, diss, bc));
}

template <class S>
__global__ void __launch_bounds__(256)
    km_fused(const typename S::s2* __restrict__ uin,
             const uint8_t* __restrict__ mask, typename S::s2* __restrict__ uout,
             typename S::s1* __restrict__ div, typename S::s1* __restrict__ p1,
             int W, int H, float dt, float scale, int viscous, float alpha,
             float rbeta) {
  const int gx = (W + fused::TX - 1) / fused::TX;
  const int t = xcd_tile((int)blockIdx.x, gx * ((H + fused::TY - 1) / fused::TY));
  fused::fused_vort_div_tile<S>((t % gx) * fused::TX, (t / gx) * fused::TY, uin, mask, uout,
                                div, p1, W, H, dt, scale, viscous, alpha, rbeta);
}

template <int ROWS, int WAVES, int COLS, class S, bool CORR>
__global__ void __launch_bounds__(WAVES * 64)
    __attribute__((amdgpu_waves_per_eu(jt::jacobi_waves_per_eu<ROWS * COLS>())))
    km_jacobi(const typename S::s1* __restrict__ pin,
              const typename S::s1* __restrict__ b,
              const uint8_t* __restrict__ mask, typename S::s1* __restrict__ pout,
              float* __restrict__ pout32, int W, int H, int k) {
  // 1D grid of gx * gy tiles in XCD order
  const int gx = (W + 64 * COLS - 2 * k - 1) / (64 * COLS - 2 * k);
  const int gy = (H + ROWS * WAVES - 2 * k - 1) / (ROWS * WAVES - 2 * k);
  const int t = xcd_tile<XCD_RUNS>((int)blockIdx.x, gx * gy);
  jt::jacobi_tb_tile<ROWS, WAVES, COLS, S, CORR>(t % gx, t / gx, pin, b, mask, pout, W, H,
                                                 k, nullptr, 0, pout32);
}

template <class S>
__global__ void __launch_bounds__(256)
    km_subgrad(const typename S::s2* __restrict__ uin, const float* __restrict__ p,
               const uint8_t* __restrict__ mask, typename S::s2* __restrict__ uout,
               uint8_t* __restrict__ mask_next, int W, int H) {
  int x, y;
  cell_of(x, y, W, H);
  if (x >= W || y >= H) return;
  const int pos = y * W + x;
  S::st2(uout, pos, subgrad_cell<S>(uin, p, mask, x, y, W, H));
  if (mask_next) mask_next[pos] = mask[pos] & MASK_PERSIST;
}

template <class S>
__global__ void __launch_bounds__(256)
    km_advect_dye(const typename S::s1* __restrict__ qin,
                  typename S::s1* __restrict__ qout,
                  const typename S::s2* __restrict__ u,
                  const uint8_t* __restrict__ mask, int Wp, int Hp, int Wv,
                  int Hv, float dt, float speed, float diss) {
  int x, y;
  cell_of(x, y, Wp, Hp);
  if (x >= Wp || y >= Hp) return;
  S::st(qout, y * Wp + x,
        advect_dye_cell<S>(qin, u, mask, x, y, Wp, Hp, Wv, Hv, dt, speed, diss));
}

struct SplatBatchM {
  int n;
  Splat s[16];
};

template <class S>
__global__ void __launch_bounds__(256)
    km_splats(typename S::s2* __restrict__ u, typename S::s1* __restrict__ q,
              int W, int H, SplatBatchM sb) {
  int x, y;
  cell_of(x, y, W, H);
  if (x >= W || y >= H) return;
  const int pos = y * W + x;
  float2 v = S::ld2(u, pos);
  float d = q ? S::ld(q, pos) : 0.0f;
  for (int i = 0; i < sb.n; ++i) {
    const float w = gaussian_weight(sb.s[i], x, y, W, H);
    v.x = v.x + sb.s[i].ax * w;
    v.y = v.y + sb.s[i].ay * w;
    d = d + sb.s[i].dye * w;
  }
  S::st2(u, pos, v);
  if (q) S::st(q, pos, d);
}

template <class S>
hipError_t mixed_dispatch_advect(const void* uin, void* uout, const uint8_t* mask,
                                 int W, int H, float dt, float speed, float diss,
                                 BcParams bc, hipStream_t s) {
  km_advect<S><<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      (const typename S::s2*)uin, (typename S::s2*)uout, mask, W, H, dt, speed,
      diss, bc);
  return hipGetLastError();
}

template <int R, int WV, int CL, class S>
hipError_t mixed_tb(const void* pin, const void* b, const uint8_t* mask,
                    void* pout, float* pout32, int W, int H, int k, bool corr,
                    hipStream_t s) {
  constexpr int EX = 64 * CL, EY = R * WV;
  const dim3 grid(((W + EX - 2 * k - 1) / (EX - 2 * k)) * ((H + EY - 2 * k - 1) / (EY - 2 * k)),
                  1, 1);
  if (corr)
    km_jacobi<R, WV, CL, S, true><<<grid, WV * 64, 0, s>>>(
        (const typename S::s1*)pin, (const typename S::s1*)b, mask,
        (typename S::s1*)pout, pout32, W, H, k);
  else
    km_jacobi<R, WV, CL, S, false><<<grid, WV * 64, 0, s>>>(
        (const typename S::s1*)pin, (const typename S::s1*)b, mask,
        (typename S::s1*)pout, pout32, W, H, k);
  return hipGetLastError();
}

template <class S>
hipError_t mixed_tb_shape(const void* pin, const void* b, const uint8_t* mask,
                          void* pout, float* pout32, int W, int H, JacobiCfg c,
                          bool corr, hipStream_t s) {
#define NX_MTB(R, WV, CL)                                                 \
  if (c.rows == R && c.waves == WV && c.cols == CL)                       \
    return mixed_tb<R, WV, CL, S>(pin, b, mask, pout, pout32, W, H, c.k, corr, s);
  NX_MTB(4, 8, 1)
  NX_MTB(8, 8, 1)
  NX_MTB(8, 16, 1)
  NX_MTB(16, 8, 1)
  NX_MTB(2, 8, 2)
  NX_MTB(2, 16, 2)
  NX_MTB(4, 4, 2)
  NX_MTB(4, 8, 2)
  NX_MTB(4, 16, 2)
  NX_MTB(8, 4, 2)
  NX_MTB(8, 8, 2)
#undef NX_MTB
  return hipErrorInvalidValue;
}

}  // namespace

#define NX_BY_STORAGE(st, call_bf16, call_f16) \
  ((st) == ST_BF16 ? (call_bf16) : (st) == ST_F16 ? (call_f16) : hipErrorInvalidValue)

hipError_t launch_m_advect(int st, const void* uin, void* uout, const uint8_t* mask,
                           int W, int H, float dt, float speed, float diss,
                           BcParams bc, hipStream_t s) {
  return NX_BY_STORAGE(
      st, mixed_dispatch_advect<BF16S>(uin, uout, mask, W, H, dt, speed, diss, bc, s),
      mixed_dispatch_advect<F16S>(uin, uout, mask, W, H, dt, speed, diss, bc, s));
}

template <class S>
static hipError_t fused_m(const void* uin, const uint8_t* mask, void* uout, void* div,
                          void* p1, int W, int H, float dt, float scale,
                          int viscous, float alpha, float rbeta, hipStream_t s) {
  const dim3 grid(((W + fused::TX - 1) / fused::TX) * ((H + fused::TY - 1) / fused::TY), 1, 1);
  km_fused<S><<<grid, dim3(64, 4), 0, s>>>(
      (const typename S::s2*)uin, mask, (typename S::s2*)uout,
      (typename S::s1*)div, (typename S::s1*)p1, W, H, dt, scale, viscous, alpha,
      rbeta);
  return hipGetLastError();
}

hipError_t launch_m_fused(int st, const void* uin, const uint8_t* mask, void* uout,
                          void* div, void* p1, int W, int H, float dt, float scale,
                          int viscous, float alpha, float rbeta, hipStream_t s) {
  return NX_BY_STORAGE(st,
                       fused_m<BF16S>(uin, mask, uout, div, p1, W, H, dt, scale,
                                      viscous, alpha, rbeta, s),
                       fused_m<F16S>(uin, mask, uout, div, p1, W, H, dt, scale,
                                     viscous, alpha, rbeta, s));
}

hipError_t launch_m_jacobi(int st, const void* pin, const void* b, const uint8_t* mask,
                           void* pout, float* pout32, int W, int H, JacobiCfg c,
                           bool corr, hipStream_t s) {
  if (!jacobi_cfg_supported(c)) return hipErrorInvalidValue;
  return NX_BY_STORAGE(
      st, mixed_tb_shape<BF16S>(pin, b, mask, pout, pout32, W, H, c, corr, s),
      mixed_tb_shape<F16S>(pin, b, mask, pout, pout32, W, H, c, corr, s));
}

template <class S>
static hipError_t subgrad_m(const void* uin, const float* p, const uint8_t* mask,
                            void* uout, uint8_t* mask_next, int W, int H,
                            hipStream_t s) {
  km_subgrad<S><<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
      (const typename S::s2*)uin, p, mask, (typename S::s2*)uout, mask_next, W, H);
  return hipGetLastError();
}

hipError_t launch_m_subgrad(int st, const void* uin, const float* p,
                            const uint8_t* mask, void* uout, uint8_t* mask_next,
                            int W, int H, hipStream_t s) {
  return NX_BY_STORAGE(st, subgrad_m<BF16S>(uin, p, mask, uout, mask_next, W, H, s),
                       subgrad_m<F16S>(uin, p, mask, uout, mask_next, W, H, s));
}

template <class S>
static hipError_t dye_m(const void* qin, void* qout, const void* u,
                        const uint8_t* mask, int Wp, int Hp, int Wv, int Hv,
                        float dt, float speed, float diss, hipStream_t s) {
  km_advect_dye<S><<<grid2d(Wp, Hp), dim3(BX, BY), 0, s>>>(
      (const typename S::s1*)qin, (typename S::s1*)qout, (const typename S::s2*)u,
      mask, Wp, Hp, Wv, Hv, dt, speed, diss);
  return hipGetLastError();
}

hipError_t launch_m_advect_dye(int st, const void* qin, void* qout, const void* u,
                               const uint8_t* mask, int Wp, int Hp, int Wv, int Hv,
                               float dt, float speed, float diss, hipStream_t s) {
  return NX_BY_STORAGE(
      st, dye_m<BF16S>(qin, qout, u, mask, Wp, Hp, Wv, Hv, dt, speed, diss, s),
      dye_m<F16S>(qin, qout, u, mask, Wp, Hp, Wv, Hv, dt, speed, diss, s));
}

template <class S>
static hipError_t splats_m(void* u, void* q, int W, int H, const Splat* sp, int n,
                           hipStream_t s) {
  for (int i0 = 0; i0 < n; i0 += 16) {
    SplatBatchM sb;
    sb.n = n - i0 < 16 ? n - i0 : 16;
    for (int i = 0; i < sb.n; ++i) sb.s[i] = sp[i0 + i];
    km_splats<S><<<grid2d(W, H), dim3(BX, BY), 0, s>>>(
        (typename S::s2*)u, (typename S::s1*)q, W, H, sb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_m_splats(int st, void* u, void* q, int W, int H, const Splat* sp,
                           int n, hipStream_t s) {
  return NX_BY_STORAGE(st, splats_m<BF16S>(u, q, W, H, sp, n, s),
                       splats_m<F16S>(u, q, W, H, sp, n, s));
}

}  // namespace nx
