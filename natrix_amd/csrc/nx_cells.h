// Per-cell device functions shared by the single-grid and batched kernels,
// so both evaluate the reference shaders with identical code.
#pragma once
#include "nx_common.h"

namespace nx {

// S2, shader.AdvectVelocity.comp:27-49, velocity read through the wall
// treatment of S1
template <class S = F32S>
__device__ __forceinline__ float2 advect_cell(
    const typename S::s2* __restrict__ uin, const uint8_t* __restrict__ mask,
    int x, int y, int W, int H, float dt, float speed, float diss,
    const BcParams& bc) {
  if (mask[y * W + x] != 0) return make_float2(0.0f, 0.0f);
  const float2 vel = load_bc<S>(uin, x, y, W, H, bc);
  const float fx = (float)x - (vel.x * dt) * speed;
  const float fy = (float)y - (vel.y * dt) * speed;
  const Bilin c = bilinear_corners(fx, fy, W, H);
  const float2 v = bilinear_mix<float2>(
      load_bc<S>(uin, c.bx, c.ty, W, H, bc), load_bc<S>(uin, c.tx, c.ty, W, H, bc),
      load_bc<S>(uin, c.bx, c.by, W, H, bc), load_bc<S>(uin, c.tx, c.by, W, H, bc),
      c.dx, c.dy);
  return make_float2(clampf(v.x * diss, -1.0f, 1.0f),
                     clampf(v.y * diss, -1.0f, 1.0f));
}

// S9, shader.SubtractGradient.comp:24-46 (pressure always float32)
template <class S = F32S>
__device__ __forceinline__ float2 subgrad_cell(
    const typename S::s2* __restrict__ uin, const float* __restrict__ p,
    const uint8_t* __restrict__ mask, int x, int y, int W, int H) {
  const int pos = y * W + x;
  const int iL = y * W + max(x - 1, 0), iR = y * W + min(x + 1, W - 1);
  const int iB = max(y - 1, 0) * W + x, iT = min(y + 1, H - 1) * W + x;
  const float pc = p[pos];
  const float x1 = mask[iL] ? pc : p[iL];
  const float x2 = mask[iR] ? pc : p[iR];
  const float y1 = mask[iB] ? pc : p[iB];
  const float y2 = mask[iT] ? pc : p[iT];
  float2 v = S::ld2(uin, pos);
  v.x -= 0.5f * (x2 - x1);
  v.y -= 0.5f * (y2 - y1);
  return v;
}

// S11, demo/shaders/shader.AdvectParticle.comp:21-69 (dye grid P, velocity
// grid V)
template <class S = F32S>
__device__ __forceinline__ float advect_dye_cell(
    const typename S::s1* __restrict__ qin, const typename S::s2* __restrict__ u,
    const uint8_t* __restrict__ mask, int x, int y, int Wp, int Hp, int Wv,
    int Hv, float dt, float speed, float diss) {
  const float nx = ((float)x / (float)Wp) * (float)Wv;
  const float ny = ((float)y / (float)Hp) * (float)Hv;
  // uint() truncation (n >= 0); x/Wp < 1 keeps it below Wv, the min() only
  // guards the memory access
  const int opos = min((int)ny, Hv - 1) * Wv + min((int)nx, Wv - 1);
  if (mask[opos] != 0) return 0.0f;
  const Bilin cv = bilinear_corners(nx, ny, Wv, Hv);
  const float2 vel =
      bilinear_mix<float2>(S::ld2(u, cv.lt), S::ld2(u, cv.rt), S::ld2(u, cv.lb),
                           S::ld2(u, cv.rb), cv.dx, cv.dy);
  const float sx = (float)Wp / (float)Wv, sy = (float)Hp / (float)Hv;
  const float vx = vel.x * sx, vy = vel.y * sy;
  const float fx = (float)x - (vx * dt) * speed;
  const float fy = (float)y - (vy * dt) * speed;
  const Bilin c = bilinear_corners(fx, fy, Wp, Hp);
  return bilinear_mix<float>(S::ld(qin, c.lt), S::ld(qin, c.rt),
                             S::ld(qin, c.lb), S::ld(qin, c.rb), c.dx, c.dy) *
         diss;
}

// e2: w = float(exp(double(-(dx*dx + dy*dy) / (r*r))))
__device__ __forceinline__ float gaussian_weight(const Splat& s, int x, int y,
                                                 int W, int H) {
  const float sx = s.px * (float)W, sy = s.py * (float)H;
  const float dx = sx - (float)x, dy = sy - (float)y;
  const float e = -(dx * dx + dy * dy) / (s.radius * s.radius);
  return (float)exp((double)e);
}

}  // namespace nx
