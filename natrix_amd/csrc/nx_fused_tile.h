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

// Stages of one 64x16 output tile with halo HA (3, or 4 with viscosity);
// every stage walks its region as one flat index range over the block's 256
// threads, so no lane idles on the few columns past a multiple of 64.  In
// an INTERIOR tile (halo inside the grid) no coordinate clamps, and the
// stencil neighbours of a stage slot are fixed offsets in the previous
// stage's array; border tiles evaluate the clamped coordinates.
template <class S, bool VISC, bool INTERIOR>
__device__ __forceinline__ void fused_tile_stages(
    float2* A, float* Wc, float2* V2, const Tile& t, const typename S::s2* __restrict__ uin,
    const uint8_t* __restrict__ mask, typename S::s2* __restrict__ uout,
    typename S::s1* __restrict__ div, typename S::s1* __restrict__ p1, float dt,
    float scale, float alpha, float rbeta) {
  constexpr int hA = VISC ? 4 : 3, hW = hA - 1, hV2 = hA - 2;
  constexpr int WA = TX + 2 * hA, HA = TY + 2 * hA;
  constexpr int WW = TX + 2 * hW, HW = TY + 2 * hW;
  constexpr int WV = TX + 2 * hV2, HV = TY + 2 * hV2;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int W = t.W;
  // the viscosity output reuses A's storage: A is last read by the
  // confinement stage, which a barrier separates from the viscosity stage
  float2* const V3 = A;

  // 1. advected velocity at clamped coordinates
  for (int k = tid; k < WA * HA; k += 256) {
    const int j = k / WA, i = k - j * WA;
    if (INTERIOR) {
      A[k] = S::ld2(uin, (t.oy - hA + j) * W + (t.ox - hA + i));
    } else {
      const int gy = t.cy(t.oy - hA + j), gx = t.cx(t.ox - hA + i);
      A[k] = S::ld2(uin, gy * W + gx);
    }
  }
  __syncthreads();

  // 2. curl, shader.CalcVorticity.comp:21-26
  for (int k = tid; k < WW * HW; k += 256) {
    const int j = k / WW, i = k - j * WW;
    float2 vL, vR, vB, vT;
    if (INTERIOR) {
      const int a = (j + 1) * WA + (i + 1);
      vL = A[a - 1];
      vR = A[a + 1];
      vB = A[a - WA];
      vT = A[a + WA];
    } else {
      const int gy = t.cy(t.oy - hW + j), gx = t.cx(t.ox - hW + i);
      const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
      const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
      vL = A[sidx(t, hA, xl, gy)];
      vR = A[sidx(t, hA, xr, gy)];
      vB = A[sidx(t, hA, gx, yb)];
      vT = A[sidx(t, hA, gx, yt)];
    }
    Wc[k] = 0.5f * ((vR.y - vL.y) - (vT.x - vB.x));
  }
  __syncthreads();

  // 3. confinement, shader.ApplyVorticity.comp:26-39
  for (int k = tid; k < WV * HV; k += 256) {
    const int j = k / WV, i = k - j * WV;
    if (INTERIOR) {
      const int w = (j + 1) * WW + (i + 1);
      V2[k] = confine_cell(A[(j + 2) * WA + (i + 2)], Wc[w], Wc[w - 1], Wc[w + 1],
                           Wc[w - WW], Wc[w + WW], dt, scale);
    } else {
      const int gy = t.cy(t.oy - hV2 + j), gx = t.cx(t.ox - hV2 + i);
      const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
      const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
      V2[k] = confine_cell(A[sidx(t, hA, gx, gy)], Wc[sidx(t, hW, gx, gy)],
                           Wc[sidx(t, hW, xl, gy)], Wc[sidx(t, hW, xr, gy)],
                           Wc[sidx(t, hW, gx, yb)], Wc[sidx(t, hW, gx, yt)], dt, scale);
    }
  }
  __syncthreads();

  // 4. viscosity, shader.Viscosity.comp:25-31 (only when nu > 0)
  const float2* Vf = V2;
  int hf = hV2;
  if (VISC) {
    constexpr int W3 = TX + 2, H3 = TY + 2;
    for (int k = tid; k < W3 * H3; k += 256) {
      const int j = k / W3, i = k - j * W3;
      float2 L, R, B, T, c;
      if (INTERIOR) {
        const int v = (j + 1) * WV + (i + 1);
        L = V2[v - 1];
        R = V2[v + 1];
        B = V2[v - WV];
        T = V2[v + WV];
        c = V2[v];
      } else {
        const int gy = t.cy(t.oy - 1 + j), gx = t.cx(t.ox - 1 + i);
        const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
        const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
        L = V2[sidx(t, hV2, xl, gy)];
        R = V2[sidx(t, hV2, xr, gy)];
        B = V2[sidx(t, hV2, gx, yb)];
        T = V2[sidx(t, hV2, gx, yt)];
        c = V2[sidx(t, hV2, gx, gy)];
      }
      V3[k] = make_float2(((((L.x + R.x) + B.x) + T.x) + c.x * alpha) * rbeta,
                          ((((L.y + R.y) + B.y) + T.y) + c.y * alpha) * rbeta);
    }
    __syncthreads();
    Vf = V3;
    hf = 1;
  }

  // 5. divergence (Divergence.comp:22-40) and the first sweep from p = 0:
  //    ((((0+0)+0)+0) - b) * 0.25 = (0 - b) * 0.25 (Poisson.comp:37)
  const int Wf = TX + 2 * hf;
  for (int k = tid; k < TX * TY; k += 256) {
    const int j = k / TX, i = k - j * TX;
    const int gy = t.oy + j, gx = t.ox + i;
    if (!INTERIOR && (gy >= t.H || gx >= W)) continue;
    const int pos = gy * W + gx;
    float2 vl, vr, vb, vt, vc;
    int ml, mr, mb, mt;
    if (INTERIOR) {
      const int v = (j + hf) * Wf + (i + hf);
      vl = Vf[v - 1];
      vr = Vf[v + 1];
      vb = Vf[v - Wf];
      vt = Vf[v + Wf];
      vc = Vf[v];
      ml = pos - 1;
      mr = pos + 1;
      mb = pos - W;
      mt = pos + W;
    } else {
      const int yb = t.cy(gy - 1), yt = t.cy(gy + 1);
      const int xl = t.cx(gx - 1), xr = t.cx(gx + 1);
      vl = Vf[sidx(t, hf, xl, gy)];
      vr = Vf[sidx(t, hf, xr, gy)];
      vb = Vf[sidx(t, hf, gx, yb)];
      vt = Vf[sidx(t, hf, gx, yt)];
      vc = Vf[sidx(t, hf, gx, gy)];
      ml = gy * W + xl;
      mr = gy * W + xr;
      mb = yb * W + gx;
      mt = yt * W + gx;
    }
    const float x1 = mask[ml] ? 0.0f : vl.x;
    const float x2 = mask[mr] ? 0.0f : vr.x;
    const float y1 = mask[mb] ? 0.0f : vb.y;
    const float y2 = mask[mt] ? 0.0f : vt.y;
    const float d = 0.5f * ((x2 - x1) + (y2 - y1));
    S::st2(uout, pos, vc);
    S::st(div, pos, d);
    if (p1) S::st(p1, pos, (0.0f - d) * 0.25f);
  }
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
  const Tile t{ox, oy, W, H};
  const int h = viscous ? 4 : 3;
  const bool interior = ox - h >= 0 && oy - h >= 0 && ox + TX + h <= W && oy + TY + h <= H;
  if (viscous) {
    if (interior)
      fused_tile_stages<S, true, true>(A, Wc, V2, t, uin, mask, uout, div, p1, dt, scale,
                                       alpha, rbeta);
    else
      fused_tile_stages<S, true, false>(A, Wc, V2, t, uin, mask, uout, div, p1, dt, scale,
                                        alpha, rbeta);
  } else {
    if (interior)
      fused_tile_stages<S, false, true>(A, Wc, V2, t, uin, mask, uout, div, p1, dt, scale,
                                        alpha, rbeta);
    else
      fused_tile_stages<S, false, false>(A, Wc, V2, t, uin, mask, uout, div, p1, dt, scale,
                                         alpha, rbeta);
  }
}


}  // namespace fused
}  // namespace nx
