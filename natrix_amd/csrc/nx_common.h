// Shared device helpers for the Natrix fluid-step kernels (gfx950).
//
// Arithmetic follows the reference shaders expression by expression; the
// library is compiled with -ffp-contract=off so every + - * / rounds once,
// as a float32 shader evaluation without FMA contraction does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nx {

// u8 obstacle mask bits.  The reference stores obstacles as float2
// (natrix/core/fluid_simulator.py:368) and tests `x > 0 || y > 0`
// (shader.AdvectVelocity.comp:30); channel x <- bit0, channel y <- bit1.
// bit2 is the persistent obstacle map, which survives the end-of-step clear.
constexpr uint8_t MASK_DYN_X = 1;
constexpr uint8_t MASK_DYN_Y = 2;
constexpr uint8_t MASK_PERSIST = 4;

// shader.ApplyVorticity.comp:34
constexpr float VORTICITY_EPSILON = 2.4414e-4f;

enum BcMode { BC_REFERENCE = 0, BC_FREE_SLIP = 1, BC_INFLOW = 2 };

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// GLSL clamp(x, lo, hi) = min(max(x, lo), hi)
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
  return fminf(fmaxf(v, lo), hi);
}

// GLSL mix as x*(1-a) + y*a.
__device__ __forceinline__ float mixf(float a, float b, float t) {
  return a * (1.0f - t) + b * t;
}

__device__ __forceinline__ float2 mix2(float2 a, float2 b, float t) {
  return make_float2(mixf(a.x, b.x, t), mixf(a.y, b.y, t));
}

// Corner indices and weights of the clamped-corner bilinear fetch
// (shader.AdvectVelocity.comp:38-48, AdvectParticle.comp:23-34 and 58-69).
struct Bilin {
  int lt, rt, lb, rb;  // row-major indices
  int tx, ty, bx, by;  // clamped corner coordinates
  float dx, dy;
};

__device__ __forceinline__ Bilin bilinear_corners(float fx, float fy, int W,
                                                  int H) {
  const float mx = (float)(W - 1), my = (float)(H - 1);
  const int tx = (int)clampf(ceilf(fx), 0.0f, mx);
  const int ty = (int)clampf(ceilf(fy), 0.0f, my);
  const int bx = (int)clampf(floorf(fx), 0.0f, mx);
  const int by = (int)clampf(floorf(fy), 0.0f, my);
  Bilin r;
  r.dx = fx - (float)bx;
  r.dy = fy - (float)by;
  r.lt = ty * W + bx;
  r.rt = ty * W + tx;
  r.lb = by * W + bx;
  r.rb = by * W + tx;
  r.tx = tx;
  r.ty = ty;
  r.bx = bx;
  r.by = by;
  return r;
}

template <typename T>
__device__ __forceinline__ T bilinear_mix(T lt, T rt, T lb, T rb, float dx,
                                          float dy);

template <>
__device__ __forceinline__ float bilinear_mix<float>(float lt, float rt,
                                                     float lb, float rb,
                                                     float dx, float dy) {
  const float h1 = mixf(lt, rt, dx);
  const float h2 = mixf(lb, rb, dx);
  return mixf(h2, h1, dy);
}

template <>
__device__ __forceinline__ float2 bilinear_mix<float2>(float2 lt, float2 rt,
                                                       float2 lb, float2 rb,
                                                       float dx, float dy) {
  const float2 h1 = mix2(lt, rt, dx);
  const float2 h2 = mix2(lb, rb, dx);
  return mix2(h2, h1, dy);
}

// Boundary treatment evaluated on read.  The reference zeroes the ring in
// place before advection (shader.InitBoundaries.comp:18-34); the extra modes
// (free-slip, left inflow with right outflow) are the build's own and are
// specified in oracle/fluid.py:enforce_boundaries.  Every ring value depends
// only on the cell itself and interior cells, so evaluating it on read is
// identical to the in-place pass.
struct BcParams {
  int mode;
  int has_borders;
  float inflow;
};

__device__ __forceinline__ float2 load_bc(const float2* __restrict__ u, int x,
                                          int y, int W, int H,
                                          const BcParams& bc) {
  const float2 v = u[y * W + x];
  const bool left = x == 0, right = x == W - 1, bot = y == 0, top = y == H - 1;
  if (!(left || right || bot || top)) return v;
  if (bc.mode == BC_REFERENCE) {
    return bc.has_borders ? make_float2(0.0f, 0.0f) : v;
  }
  if (bc.mode == BC_FREE_SLIP) {
    if ((left || right) && (bot || top)) return make_float2(0.0f, 0.0f);
    if (left) return make_float2(0.0f, u[y * W + 1].y);
    if (right) return make_float2(0.0f, u[y * W + W - 2].y);
    if (bot) return make_float2(u[W + x].x, 0.0f);
    return make_float2(u[(H - 2) * W + x].x, 0.0f);
  }
  // BC_INFLOW
  if (left) return make_float2(bc.inflow, 0.0f);
  if (right) {
    if (bot || top) return make_float2(0.0f, 0.0f);
    return u[y * W + W - 2];
  }
  if (bot) return make_float2(u[W + x].x, 0.0f);
  return make_float2(u[(H - 2) * W + x].x, 0.0f);
}

struct Splat {  // Gaussian splat: u += (ax, ay) * w, dye += dye * w
  float px, py, ax, ay, radius, dye;
};

}  // namespace nx
