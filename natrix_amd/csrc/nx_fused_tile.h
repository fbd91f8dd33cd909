// Fused curl -> confinement -> [viscosity] -> divergence -> first sweep,
// one LDS-staged 64x16 tile (shared by the single-grid and batched kernels).
#pragma once
#include "nx_common.h"

namespace nx {
namespace fused {

constexpr int TX = 64, TY = 16;  // output tile
constexpr int HMAX = 4;

struct Tile {
  int ox, oy, W, H;
  __device__ int cx(int vx) const { return clampi(vx, 0, W - 1); }
  __device__ int cy(int vy) const { return clampi(vy, 0, H - 1); }
};

// index of grid coordinate (gx, gy) in a stage array of halo hs
__device__ __forceinline__ int sidx(const Tile& t, int hs, int gx, int gy) {
  return (gy - t.oy + hs) * (TX + 2 * hs) + (gx - t.ox + hs);
}

__device__ __forceinline__ float2 confine_cell(float2 u, float wc, float wl,
                                               float wr, float wb, float wt,
                                               float dt, float scale) {
  // shader.ApplyVorticity.comp:33-39
  float fx = 0.5f * (fabsf(wt) - fabsf(wb));
  float fy = 0.5f * (fabsf(wr) - fabsf(wl));
  const float mag = fmaxf(VORTICITY_EPSILON, fx * fx + fy * fy);
  const float inv = 1.0f / sqrtf(mag);
  fx = fx * inv;
  fy = fy * inv;
  const float s = scale * wc;
  fx = fx * s;
  fy = fy * (-s);
  return make_float2(u.x + fx * dt, u.y + fy * dt);
}

// One 64x16 output tile with origin (ox, oy); 256 threads as (64, 4).
// Velocity, divergence and p1 are stored in format S; the chain is float32.
template <class S = F32S>
__device__ __forceinline__ void fused_vort_div_tile(
    int ox, int oy, const typename S::s2* __restrict__ uin,
    const uint8_t* __restrict__ mask, typename S::s2* __restrict__ uout,
    typename S::s1* __restrict__ div, typename S::s1* __restrict__ p1, int W,
    int H, float dt, float scale, int viscous, float alpha, float rbeta) {
  __shared__ float2 A[(TX + 2 * HMAX) * (TY + 2 * HMAX)];
  __shared__ float Wc[(TX + 2 * HMAX - 2) * (TY + 2 * HMAX - 2)];
  __shared__ float2 V2[(TX + 2 * HMAX - 4) * (TY + 2 * HMAX - 4)];
  // the viscosity output reuses A's storage: A is last read by the
  // confinement stage, which a barrier separates from the viscosity stage
  float2* const V3 = A;

  const Tile t{ox, oy, W, H};
  const int hA = 3 + (viscous ? 1 : 0), hW = hA - 1, hV2 = hA - 2;
  const int tx = threadIdx.x, ty = threadIdx.y;

  // 1. advected velocity at clamped coordinates
  for (int j = ty; j < TY + 2 * hA; j += 4) {
    const int gy = t.cy(t.oy - hA + j);
    for (int i = tx; i < TX + 2 * hA; i += 64) {
      const int gx = t.cx(t.ox - hA + i);
      A[j * (TX + 2 * hA) + i] = S::ld2(uin, gy * W + gx);
    }
  }
  __syncthreads();

  // 2. curl, shader.CalcVorticity.comp:21-26
  for (int j = ty; j < TY + 2 * hW; j += 4) {
    const int gy = t.cy(t.oy - hW + j);
    const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
    for (int i = tx; i < TX + 2 * hW; i += 64) {
      const int gx = t.cx(t.ox - hW + i);
      const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
      const float2 vL = A[sidx(t, hA, xl, gy)];
      const float2 vR = A[sidx(t, hA, xr, gy)];
      const float2 vB = A[sidx(t, hA, gx, yb)];
      const float2 vT = A[sidx(t, hA, gx, yt)];
      Wc[j * (TX + 2 * hW) + i] = 0.5f * ((vR.y - vL.y) - (vT.x - vB.x));
    }
  }
  __syncthreads();

  // 3. confinement, shader.ApplyVorticity.comp:26-39
  for (int j = ty; j < TY + 2 * hV2; j += 4) {
    const int gy = t.cy(t.oy - hV2 + j);
    const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
    for (int i = tx; i < TX + 2 * hV2; i += 64) {
      const int gx = t.cx(t.ox - hV2 + i);
      const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
      V2[j * (TX + 2 * hV2) + i] = confine_cell(
          A[sidx(t, hA, gx, gy)], Wc[sidx(t, hW, gx, gy)],
          Wc[sidx(t, hW, xl, gy)], Wc[sidx(t, hW, xr, gy)],
          Wc[sidx(t, hW, gx, yb)], Wc[sidx(t, hW, gx, yt)], dt, scale);
    }
  }
  __syncthreads();

  // 4. viscosity, shader.Viscosity.comp:25-31 (only when nu > 0)
  const float2* Vf = V2;
  int hf = hV2;
  if (viscous) {
    for (int j = ty; j < TY + 2; j += 4) {
      const int gy = t.cy(t.oy - 1 + j);
      const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
      for (int i = tx; i < TX + 2; i += 64) {
        const int gx = t.cx(t.ox - 1 + i);
        const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
        const float2 L = V2[sidx(t, hV2, xl, gy)];
        const float2 R = V2[sidx(t, hV2, xr, gy)];
        const float2 B = V2[sidx(t, hV2, gx, yb)];
        const float2 T = V2[sidx(t, hV2, gx, yt)];
        const float2 c = V2[sidx(t, hV2, gx, gy)];
        V3[j * (TX + 2) + i] =
            make_float2(((((L.x + R.x) + B.x) + T.x) + c.x * alpha) * rbeta,
                        ((((L.y + R.y) + B.y) + T.y) + c.y * alpha) * rbeta);
      }
    }
    __syncthreads();
    Vf = V3;
    hf = 1;
  }

  // 5. divergence (Divergence.comp:22-40) and the first sweep from p = 0:
  //    ((((0+0)+0)+0) - b) * 0.25 = (0 - b) * 0.25 (Poisson.comp:37)
  for (int j = ty; j < TY; j += 4) {
    const int gy = t.oy + j;
    if (gy >= H) break;
    const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
    for (int i = tx; i < TX; i += 64) {
      const int gx = t.ox + i;
      if (gx >= W) break;
      const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
      const float x1 = mask[gy * W + xl] ? 0.0f : Vf[sidx(t, hf, xl, gy)].x;
      const float x2 = mask[gy * W + xr] ? 0.0f : Vf[sidx(t, hf, xr, gy)].x;
      const float y1 = mask[yb * W + gx] ? 0.0f : Vf[sidx(t, hf, gx, yb)].y;
      const float y2 = mask[yt * W + gx] ? 0.0f : Vf[sidx(t, hf, gx, yt)].y;
      const float d = 0.5f * ((x2 - x1) + (y2 - y1));
      const int pos = gy * W + gx;
      S::st2(uout, pos, Vf[sidx(t, hf, gx, gy)]);
      S::st(div, pos, d);
      if (p1) S::st(p1, pos, (0.0f - d) * 0.25f);
    }
  }
}


}  // namespace fused
}  // namespace nx
