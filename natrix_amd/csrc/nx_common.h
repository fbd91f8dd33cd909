// Shared device helpers for the Natrix fluid-step kernels (gfx950).
//
// Arithmetic follows the reference shaders expression by expression; the
// library is compiled with -ffp-contract=off so every + - * / rounds once,
// as a float32 shader evaluation without FMA contraction does.
#pragma once
#include <hip/hip_fp16.h>
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

// ---- storage formats --------------------------------------------------
// Velocity, pressure, divergence and dye may be stored as float32, bfloat16
// or float16; every kernel widens on load and computes in float32, and
// rounds to nearest even on store.  Velocity pairs are stored interleaved
// (x, y) in one 8-, 4- or 4-byte word.
enum Storage { ST_F32 = 0, ST_BF16 = 1, ST_F16 = 2 };

struct F32S {
  using s1 = float;
  using s2 = float2;
  static constexpr int id = ST_F32;
  __device__ static __forceinline__ float ld(const s1* p, long long i) { return p[i]; }
  __device__ static __forceinline__ void st(s1* p, long long i, float v) { p[i] = v; }
  __device__ static __forceinline__ float2 ld2(const s2* p, long long i) { return p[i]; }
  __device__ static __forceinline__ void st2(s2* p, long long i, float2 v) { p[i] = v; }
  // two adjacent scalars p[i], p[i+1] (i even)
  __device__ static __forceinline__ float2 ldp(const s1* p, long long i) {
    return *reinterpret_cast<const float2*>(p + i);
  }
  __device__ static __forceinline__ void stp(s1* p, long long i, float2 v) {
    *reinterpret_cast<float2*>(p + i) = v;
  }
};

__device__ __forceinline__ float bf16_to_f32(uint32_t h) {
  return __uint_as_float(h << 16);
}

// round to nearest, ties to even (finite inputs)
__device__ __forceinline__ uint32_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

struct BF16S {
  using s1 = uint16_t;  // bfloat16 bits
  using s2 = uint32_t;  // x in bits 0..15, y in bits 16..31
  static constexpr int id = ST_BF16;
  __device__ static __forceinline__ float ld(const s1* p, long long i) {
    return bf16_to_f32(p[i]);
  }
  __device__ static __forceinline__ void st(s1* p, long long i, float v) {
    p[i] = (uint16_t)f32_to_bf16(v);
  }
  __device__ static __forceinline__ float2 ld2(const s2* p, long long i) {
    const uint32_t v = p[i];
    return make_float2(bf16_to_f32(v & 0xFFFFu), __uint_as_float(v & 0xFFFF0000u));
  }
  __device__ static __forceinline__ void st2(s2* p, long long i, float2 v) {
    p[i] = f32_to_bf16(v.x) | (f32_to_bf16(v.y) << 16);
  }
  __device__ static __forceinline__ float2 ldp(const s1* p, long long i) {
    return ld2(reinterpret_cast<const s2*>(p + i), 0);
  }
  __device__ static __forceinline__ void stp(s1* p, long long i, float2 v) {
    st2(reinterpret_cast<s2*>(p + i), 0, v);
  }
};

struct F16S {
  using s1 = __half;
  using s2 = __half2;
  static constexpr int id = ST_F16;
  __device__ static __forceinline__ float ld(const s1* p, long long i) {
    return __half2float(p[i]);
  }
  __device__ static __forceinline__ void st(s1* p, long long i, float v) {
    p[i] = __float2half_rn(v);
  }
  __device__ static __forceinline__ float2 ld2(const s2* p, long long i) {
    return __half22float2(p[i]);
  }
  __device__ static __forceinline__ void st2(s2* p, long long i, float2 v) {
    p[i] = __floats2half2_rn(v.x, v.y);
  }
  __device__ static __forceinline__ float2 ldp(const s1* p, long long i) {
    return ld2(reinterpret_cast<const s2*>(p + i), 0);
  }
  __device__ static __forceinline__ void stp(s1* p, long long i, float2 v) {
    st2(reinterpret_cast<s2*>(p + i), 0, v);
  }
};

template <class S>
__device__ __forceinline__ float2 load_bc(const typename S::s2* __restrict__ u,
                                          int x, int y, int W, int H,
                                          const BcParams& bc) {
  const float2 v = S::ld2(u, y * W + x);
  const bool left = x == 0, right = x == W - 1, bot = y == 0, top = y == H - 1;
  if (!(left || right || bot || top)) return v;
  if (bc.mode == BC_REFERENCE) {
    return bc.has_borders ? make_float2(0.0f, 0.0f) : v;
  }
  if (bc.mode == BC_FREE_SLIP) {
    if ((left || right) && (bot || top)) return make_float2(0.0f, 0.0f);
    if (left) return make_float2(0.0f, S::ld2(u, y * W + 1).y);
    if (right) return make_float2(0.0f, S::ld2(u, y * W + W - 2).y);
    if (bot) return make_float2(S::ld2(u, W + x).x, 0.0f);
    return make_float2(S::ld2(u, (H - 2) * W + x).x, 0.0f);
  }
  // BC_INFLOW
  if (left) return make_float2(bc.inflow, 0.0f);
  if (right) {
    if (bot || top) return make_float2(0.0f, 0.0f);
    return S::ld2(u, y * W + W - 2);
  }
  if (bot) return make_float2(S::ld2(u, W + x).x, 0.0f);
  return make_float2(S::ld2(u, (H - 2) * W + x).x, 0.0f);
}

__device__ __forceinline__ float2 load_bc(const float2* __restrict__ u, int x,
                                          int y, int W, int H,
                                          const BcParams& bc) {
  return load_bc<F32S>(u, x, y, W, H, bc);
}

struct Splat {  // Gaussian splat: u += (ax, ay) * w, dye += dye * w
  float px, py, ax, ay, radius, dye;
};

// The chip dispatches the workgroups of a launch to its 8 XCDs round-robin
// (workgroup i to XCD i mod 8), and each XCD has its own L2.  Kernels whose
// blocks read overlapping halos launch a 1D grid of T blocks and take block
// (tile) xcd_tile<M>(i, T) in workgroup i: the row-major tiles are cut into
// 8 M runs of nearly equal length (the first T mod 8 M one tile longer) and
// XCD x works through runs x, x + 8, x + 16, ... in order.  Neighbouring
// tiles of a run then run on one XCD close together in time, and the rows
// and columns they share are read from that XCD's L2 instead of from
// memory; M > 1 spreads each XCD's tiles over the whole grid, so that a
// region of costlier tiles (obstacle boundaries) is shared by all XCDs
// instead of delaying one.  Each XCD gets as many tiles as the dispatch
// gives it workgroups, so this is a bijection of [0, T) for every T and M.
// M = 1 takes shifts only; M > 1 (a power of two) two integer divisions
// per workgroup, which the many-workgroup stage kernels leave out (their
// cost showed at 65536 workgroups of 256 threads), while the tile kernels
// use XCD_RUNS = 4, timed fastest on the 2048^2 Chebyshev tile (41.3 us per
// launch against 42.8 for M = 1 and 45.9 for plain round-robin order; 2
// and 8 within 0.1 us of 4).
constexpr int XCD_RUNS = 4;
template <int M = 1>
__device__ __forceinline__ int xcd_tile(int i, int T) {
  static_assert(M >= 1 && (M & (M - 1)) == 0, "runs per XCD must be a power of two");
  const int x = i & 7, j = i >> 3;
  if (M == 1) {
    const int q = T >> 3, r = T & 7;
    return x * q + min(x, r) + j;
  }
  const int q = T / (8 * M), rem = T & (8 * M - 1);
  // runs x + 8 k with k < a have q + 1 tiles, the later ones q
  const int a = min(M, max(0, (rem - x + 7) >> 3));
  int k, off;
  if (j < a * (q + 1)) {
    k = j / (q + 1);
    off = j - k * (q + 1);
  } else {
    const int jj = j - a * (q + 1);
    k = a + (q > 0 ? jj / q : 0);
    off = q > 0 ? jj - (k - a) * q : 0;
  }
  const int r = x + 8 * k;
  return r * q + min(r, rem) + off;
}

// A 1D grid of the 64 x 4 cell blocks that cover a W x H grid (one
// wavefront per 64 consecutive x of a row), and the cell of the calling
// thread when workgroup i takes block xcd_tile(i, blocks): the rows a
// stencil shares with the blocks above and below are then read once per XCD.
inline dim3 cell_grid(int W, int H) {
  return dim3((unsigned)(((W + 63) / 64) * ((H + 3) / 4)), 1, 1);
}

__device__ __forceinline__ bool cell_of(int& x, int& y, int W, int H) {
  const int gx = (W + 63) >> 6;
  const int t = xcd_tile((int)blockIdx.x, gx * ((H + 3) >> 2));
  x = (t % gx) * 64 + (int)threadIdx.x;
  y = (t / gx) * 4 + (int)threadIdx.y;
  return x < W && y < H;
}

}  // namespace nx
