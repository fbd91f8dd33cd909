// The learned two-grid cycle on the simulator's grid (no counterpart in
// the reference): restriction of the residual to the coarse grid,
// interpolation of a coarse correction, the coarse-grid smoothers, the
// energy line search along a correction, the fine-grid smoothers of the
// cycle (Jacobi with a damped sweep; Chebyshev, nx_cheb.hip and
// nx_stream.hip), and the FP64 iterative refinement of the pressure that
// the tolerance-driven step wraps around its correction solves.  The
// operator throughout is the 5-point stencil of the Jacobi sweep
// (shader.Poisson.comp:26-37: a solid or out-of-grid neighbour reads the
// centre value).
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/natrix.h"
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"
#include "nx_sim.h"

using namespace nx;

// ---- coarse-grid transfer for the learned guess --------------------------
// Restriction: coarse cell (I, J) of factor f covers fine cells
// [fI, fI+f) x [fJ, fJ+f).  rhs_c = f^2 * mean over its fluid fine cells of
// the float32 residual r = b - A p (the correction equation in coarse grid
// units); mask_c = 1 only when every one of its fine cells is solid, so the
// coarse fluid region covers the fine one.  Two passes: per fine row y and
// coarse column I, the sum and fluid count of r over the row's f cells
// (row-major lanes, one H x Wc record each), then per coarse cell the sum of
// its f row records.
__global__ void __launch_bounds__(256)
    k_restrict_rows(const float* __restrict__ p, const float* __restrict__ b,
                    const uint8_t* __restrict__ mask, float* __restrict__ rsum,
                    float* __restrict__ rcnt, int W, int H, int f, int Wc) {
  const int X = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (X >= Wc || y >= H) return;
  float sum = 0.0f;
  int fluid = 0;
  const int x0 = X * f, x1 = min(x0 + f, W);
  const int rowB = max(y - 1, 0) * W, rowT = min(y + 1, H - 1) * W, row = y * W;
  for (int x = x0; x < x1; ++x) {
    const int pos = row + x;
    if (mask[pos]) continue;
    const int iL = row + max(x - 1, 0), iR = row + min(x + 1, W - 1);
    const int iB = rowB + x, iT = rowT + x;
    const float pc = p[pos];
    const float xl = mask[iL] ? pc : p[iL];
    const float xr = mask[iR] ? pc : p[iR];
    const float yb = mask[iB] ? pc : p[iB];
    const float yt = mask[iT] ? pc : p[iT];
    sum += b[pos] - ((((xl + xr) + yb) + yt) - 4.0f * pc);
    ++fluid;
  }
  rsum[y * Wc + X] = sum;
  rcnt[y * Wc + X] = (float)fluid;
}

// Column strips for the residual-type kernels below: wave w of a 64 x
// RES_ROWS block owns rows y0 .. y0 + RS - 1 (RS = RES_ROWS / 4) of column
// x = 64 bx + lane.  Every load of a strip is issued before any is used
// (one wait per wave, not one per row): rows y0 - 1 .. y0 + RS of the column
// (clamped to the grid, so an edge row reads the row itself, as the clamped
// index does) and, per row, the value one column beyond the lane's 64-wide
// segment (x - 1 for lane 0, x + 1 for the others, clamped).  The left and
// right neighbours then come from the adjacent lanes by DPP, lanes 0 and 63
// take the loaded values and the grid's right edge the cell itself.  Lanes
// past the last column read column W - 1 and write nothing, and rows past
// the last read row H - 1 and are skipped, so every lane runs the same code.
constexpr int RS = RES_ROWS / 4;

template <typename T, typename V>
__device__ __forceinline__ void strip_load(const T* __restrict__ a, V (&col)[RS + 2],
                                           V (&edge)[RS], int y0, int xc, int ex, int W,
                                           int H) {
#pragma unroll
  for (int i = 0; i < RS + 2; ++i) col[i] = (V)a[min(max(y0 - 1 + i, 0), H - 1) * W + xc];
#pragma unroll
  for (int i = 0; i < RS; ++i) edge[i] = (V)a[min(y0 + i, H - 1) * W + ex];
}

// neighbours of a row's cell c with the row's loaded edge value e; the
// shifts run with every lane active (no branch around them)
__device__ __forceinline__ float nbr_left(float c, float e, int lane) {
  const float s = jt::lane_from_left(c);
  return lane == 0 ? e : s;
}
__device__ __forceinline__ float nbr_right(float c, float e, int lane, bool last) {
  const float s = jt::lane_from_right(c);
  return last ? c : (lane == 63 ? e : s);
}
__device__ __forceinline__ int inbr_left(int c, int e, int lane) {
  const int s = jt::ilane_from_left(c);
  return lane == 0 ? e : s;
}
__device__ __forceinline__ int inbr_right(int c, int e, int lane, bool last) {
  const int s = jt::ilane_from_right(c);
  return last ? c : (lane == 63 ? e : s);
}

// The same row records when f is a power of two dividing 64: lane l owns
// fine column 64 b + l, and the f lanes of a coarse column add their values
// with xor shuffles inside their aligned group (lane order differs from the
// loop above only in the float32 summation order); rows stream as above.
__global__ void __launch_bounds__(256)
    k_restrict_rows_seg(const float* __restrict__ p, const float* __restrict__ b,
                        const uint8_t* __restrict__ mask, float* __restrict__ rsum,
                        float* __restrict__ rcnt, int W, int H, int f, int Wc) {
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int x = blockIdx.x * 64 + lane, xc = min(x, W - 1);
  const int ex = lane == 0 ? max(x - 1, 0) : min(x + 1, W - 1);
  const int y0 = blockIdx.y * RES_ROWS + wave * RS;
  if (y0 >= H) return;  // wave-uniform
  const bool last = x + 1 >= W;
  float pv[RS + 2], pe[RS], bv[RS];
  int mv[RS + 2], me[RS];
  strip_load(p, pv, pe, y0, xc, ex, W, H);
  strip_load(mask, mv, me, y0, xc, ex, W, H);
#pragma unroll
  for (int i = 0; i < RS; ++i) bv[i] = b[min(y0 + i, H - 1) * W + xc];
#pragma unroll
  for (int i = 0; i < RS; ++i) {
    const int y = y0 + i;
    if (y >= H) break;  // wave-uniform
    const float pc = pv[i + 1];
    const int mcv = mv[i + 1];
    const float l = nbr_left(pc, pe[i], lane), r = nbr_right(pc, pe[i], lane, last);
    const int ml = inbr_left(mcv, me[i], lane), mr = inbr_right(mcv, me[i], lane, last);
    float v = 0.0f, c = 0.0f;
    if (x < W && mcv == 0) {
      const float xl = ml ? pc : l;
      const float xr = mr ? pc : r;
      const float yb = mv[i] ? pc : pv[i];
      const float yt = mv[i + 2] ? pc : pv[i + 2];
      v = bv[i] - ((((xl + xr) + yb) + yt) - 4.0f * pc);
      c = 1.0f;
    }
    for (int off = f >> 1; off > 0; off >>= 1) {
      v += __shfl_xor(v, off, 64);
      c += __shfl_xor(c, off, 64);
    }
    if ((lane & (f - 1)) == 0 && x < W) {
      rsum[y * Wc + x / f] = v;
      rcnt[y * Wc + x / f] = c;
    }
  }
}

// Coarse cells of a padded coarse grid (Wp x Hp beyond the Wv x Hv cells
// that cover the fine grid) are solid with a zero right-hand side.  With
// `block_sum` rhs_c is the plain sum S over the block's fluid cells (f^2
// times the mean over all f x f cells, solid ones counting as zero): the
// transpose of piecewise-constant interpolation, which pairs with the
// face-weighted coarse operator of k_coarse_faces.
__global__ void __launch_bounds__(256)
    k_restrict_cols(const float* __restrict__ rsum, const float* __restrict__ rcnt,
                    float* __restrict__ rc, uint8_t* __restrict__ mc, int H, int f,
                    int Wv, int Hv, int Wp, int Hp, int block_sum) {
  const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y * 4 + threadIdx.y;
  if (X >= Wp || Y >= Hp) return;
  if (X >= Wv || Y >= Hv) {
    rc[Y * Wp + X] = 0.0f;
    mc[Y * Wp + X] = 1;
    return;
  }
  float sum = 0.0f, cnt = 0.0f;
  for (int y = Y * f; y < min(Y * f + f, H); ++y) {
    sum += rsum[y * Wv + X];
    cnt += rcnt[y * Wv + X];
  }
  rc[Y * Wp + X] = cnt == 0.0f ? 0.0f : block_sum ? sum : (float)(f * f) * (sum / cnt);
  mc[Y * Wp + X] = cnt == 0.0f ? 1 : 0;
}

// Face weights of the coarse operator that goes with the block-sum
// restriction: the east face of coarse cell (X, Y) weighs n / f, n the
// fine rows y of the block whose cells (f X + f - 1, y) and (f X + f, y)
// are both fluid and inside the grid (the open fraction of the block
// boundary); the south face likewise over the block's columns.  A face at
// the grid's edge, of a padded cell or of a solid cell weighs 0, an open
// face between full blocks 1.  The coarse operator is then
// (L e)_I = sum over faces w (e_J - e_I): a channel narrower than a block
// conducts in proportion to its width instead of as a full cell, and a
// sliver of fluid at an obstacle's edge couples only through its open
// faces.  faces[0, Hp Wp) holds the east weights, faces[Hp Wp, 2 Hp Wp)
// the south ones.
__global__ void __launch_bounds__(256)
    k_coarse_faces(const uint8_t* __restrict__ mask, int W, int H, int f, int Wp, int Hp,
                   float* __restrict__ faces) {
  const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y * 4 + threadIdx.y;
  if (X >= Wp || Y >= Hp) return;
  int ne = 0, ns = 0;
  const int xe = f * X + f, ys = f * Y + f;  // first column / row of the next block
  // every load is unconditional from a clamped in-grid address and the
  // conditions apply afterwards, so an unrolled batch of loads is in
  // flight together
  const int xa = min(xe - 1, W - 1), xb = min(xe, W - 1);
  const int ya = min(ys - 1, H - 1), yb = min(ys, H - 1);
  const bool east = xe < W, south = ys < H;
#pragma unroll 8
  for (int k = 0; k < f; ++k) {
    const int y = f * Y + k, x = f * X + k;
    const int yc = min(y, H - 1), xc = min(x, W - 1);
    const uint8_t e0 = mask[yc * W + xa], e1 = mask[yc * W + xb];
    const uint8_t s0 = mask[ya * W + xc], s1 = mask[yb * W + xc];
    ne += (east && y < H && e0 == 0 && e1 == 0) ? 1 : 0;
    ns += (south && x < W && s0 == 0 && s1 == 0) ? 1 : 0;
  }
  const int pos = Y * Wp + X;
  faces[pos] = (float)ne / (float)f;
  faces[Wp * Hp + pos] = (float)ns / (float)f;
}

// Block sums of a restricted field onto a grid f times coarser (the middle
// level of a three-level cycle onto the network's grid): rc = the sum of
// rm over the fluid cells of each f x f block of the Wm x Hm grid (rm is 0
// on its solid cells), mc = 1 when the block has none; cells of a padded
// Wc x Hc grid beyond the blocks that cover Wm x Hm are solid.
__global__ void __launch_bounds__(256)
    k_block_sum(const float* __restrict__ rm, const uint8_t* __restrict__ mm, int Wm, int Hm,
                int f, float* __restrict__ rc, uint8_t* __restrict__ mc, int Wc, int Hc) {
  const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y * 4 + threadIdx.y;
  if (X >= Wc || Y >= Hc) return;
  float s = 0.0f;
  int n = 0;
  for (int k = 0; k < f; ++k) {
    const int y = f * Y + k;
    if (y >= Hm) break;
    for (int j = 0; j < f; ++j) {
      const int x = f * X + j;
      if (x < Wm && mm[y * Wm + x] == 0) {
        s += rm[y * Wm + x];
        ++n;
      }
    }
  }
  rc[Y * Wc + X] = n ? s : 0.0f;
  mc[Y * Wc + X] = n ? 0 : 1;
}

// Prolongation: p += scale * e interpolated at the fine cell, on fluid
// cells.  Coarse cell centres sit at fine coordinate f*I + (f-1)/2
// (clamped at the edges); the four bilinear weights are renormalised over
// the fluid coarse cells among them, so solid coarse cells (which carry no
// correction) do not pull the correction toward zero next to an obstacle.
template <bool ADD>
__global__ void __launch_bounds__(256)
    k_prolong(float* __restrict__ p, const uint8_t* __restrict__ mask,
              const float* __restrict__ ec, const uint8_t* __restrict__ mc, int W, int H,
              int f, int Wc, int Hc, float scale) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= W || y >= H) return;
  const int pos = y * W + x;
  // the coarse indices depend on the cell's position only, so the mask,
  // the four coarse masks and corrections are loaded together and the
  // fine mask selects at the end (a solid cell gets no correction)
  const float cx = ((float)x - 0.5f * (float)(f - 1)) / (float)f;
  const float cy = ((float)y - 0.5f * (float)(f - 1)) / (float)f;
  const float fx = fminf(fmaxf(cx, 0.0f), (float)(Wc - 1));
  const float fy = fminf(fmaxf(cy, 0.0f), (float)(Hc - 1));
  const int x0 = (int)fx, y0 = (int)fy;
  const int x1 = min(x0 + 1, Wc - 1), y1 = min(y0 + 1, Hc - 1);
  const float tx = fx - (float)x0, ty = fy - (float)y0;
  const int i00 = y0 * Wc + x0, i01 = y0 * Wc + x1, i10 = y1 * Wc + x0, i11 = y1 * Wc + x1;
  const uint8_t m = mask[pos];
  const uint8_t m00 = mc[i00], m01 = mc[i01], m10 = mc[i10], m11 = mc[i11];
  const float e00 = ec[i00], e01 = ec[i01], e10 = ec[i10], e11 = ec[i11];
  const float pold = ADD ? p[pos] : 0.0f;
  const float w00 = m00 ? 0.0f : (1.0f - tx) * (1.0f - ty);
  const float w01 = m01 ? 0.0f : tx * (1.0f - ty);
  const float w10 = m10 ? 0.0f : (1.0f - tx) * ty;
  const float w11 = m11 ? 0.0f : tx * ty;
  const float wsum = ((w00 + w01) + w10) + w11;
  float v = 0.0f;
  if (m == 0 && wsum > 0.0f)
    v = (((w00 * e00 + w01 * e01) + w10 * e10) + w11 * e11) / wsum;
  if (ADD) {
    if (m == 0) p[pos] = pold + scale * v;
  } else {
    p[pos] = scale * v;
  }
}

// Smoothing of a coarse correction problem L e = r held entirely by one
// workgroup: lane l of wave w owns columns 2l, 2l+1 and rows 8w .. 8w+7 of a
// grid of at most 128 x 128 cells, in registers, for every iteration.
// Horizontal neighbours come from the adjacent lane by DPP, the rows next to
// a wave's strip from double-buffered LDS (one barrier per stencil).  The
// operator is the fine sweep's: a solid or out-of-grid neighbour reads the
// centre value, so -L is symmetric positive semi-definite with spectrum in
// [0, 8].  CHEB = false: Jacobi sweeps e <- e + (L e - r) / 4.  CHEB = true:
// Chebyshev iteration for -L e = -r on the interval [lmin, 8] (three-term
// recurrence of Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// Algorithm 12.1), which damps every mode above lmin by 1 / T_k((8 + lmin)
// / (8 - lmin)) after k iterations and never amplifies the modes below.
constexpr int CJ_ROWS = 8, CJ_WAVES = 16, CJ_MAX = 128;

template <int NR>
struct CoarseTile {
  float (*top)[CJ_WAVES][CJ_MAX];
  float (*bot)[CJ_WAVES][CJ_MAX];
  int code[NR][2];  // bit 0: left, 1: right, 2: up (y-1), 3: down (y+1)
  int lane, wave, x0, buf;

  // out = L v for this lane's cells (all threads of the block call it),
  // summed ((l + r) + u) + d - 4 c (the Jacobi sweep's order) or, CHEB,
  // ((u + l) + r) + d - 4 c (the Chebyshev kernels' order)
  template <bool CHEB = false>
  __device__ __forceinline__ void lap(const float (&v)[NR][2],
                                      float (&out)[NR][2]) {
    top[buf][wave][x0] = v[0][0];
    top[buf][wave][x0 + 1] = v[0][1];
    bot[buf][wave][x0] = v[NR - 1][0];
    bot[buf][wave][x0 + 1] = v[NR - 1][1];
    __syncthreads();
    float fl[NR], fr[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      fl[i] = jt::lane_from_left(v[i][1]);
      fr[i] = jt::lane_from_right(v[i][0]);
    }
    const float* above = bot[buf][wave > 0 ? wave - 1 : 0];
    const float* below = top[buf][wave < CJ_WAVES - 1 ? wave + 1 : 0];
#pragma unroll
    for (int i = 0; i < NR; ++i)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float ctr = v[i][c];
        const int cd = code[i][c];
        float L = c == 1 ? v[i][0] : fl[i];
        float R = c == 0 ? v[i][1] : fr[i];
        float U = i > 0 ? v[i - 1][c] : above[x0 + c];
        float D = i < NR - 1 ? v[i + 1][c] : below[x0 + c];
        L = (cd & 1) ? ctr : L;
        R = (cd & 2) ? ctr : R;
        U = (cd & 4) ? ctr : U;
        D = (cd & 8) ? ctr : D;
        out[i][c] = CHEB ? (((U + L) + R) + D) - 4.0f * ctr : (((L + R) + U) + D) - 4.0f * ctr;
      }
    buf ^= 1;
  }
};

template <bool CHEB>
__global__ void __launch_bounds__(1024)
    k_coarse_smooth(float* __restrict__ e, const float* __restrict__ rhs,
                    const uint8_t* __restrict__ mc, int Wc, int Hc, int iters, float lmin) {
  __shared__ float top[2][CJ_WAVES][CJ_MAX], bot[2][CJ_WAVES][CJ_MAX];
  CoarseTile<CJ_ROWS> t;
  t.top = top;
  t.bot = bot;
  t.lane = threadIdx.x & 63;
  t.wave = threadIdx.x >> 6;
  t.x0 = 2 * t.lane;
  t.buf = 0;
  const int y0 = CJ_ROWS * t.wave;
  auto solid = [&](int x, int y) {
    return x < 0 || y < 0 || x >= Wc || y >= Hc || mc[y * Wc + x] != 0;
  };
  float x[CJ_ROWS][2], r[CJ_ROWS][2], lv[CJ_ROWS][2];
#pragma unroll
  for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gx = t.x0 + c, gy = y0 + i;
      const bool in = gx < Wc && gy < Hc;
      x[i][c] = in ? e[gy * Wc + gx] : 0.0f;
      r[i][c] = in ? rhs[gy * Wc + gx] : 0.0f;
      t.code[i][c] = (solid(gx - 1, gy) ? 1 : 0) | (solid(gx + 1, gy) ? 2 : 0) |
                     (solid(gx, gy - 1) ? 4 : 0) | (solid(gx, gy + 1) ? 8 : 0);
    }
  if (!CHEB) {
    for (int s = 0; s < iters; ++s) {
      t.lap(x, lv);
#pragma unroll
      for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
        for (int c = 0; c < 2; ++c) x[i][c] = x[i][c] + (lv[i][c] - r[i][c]) * 0.25f;
    }
  } else {
    // A = -L, f = -rhs; residual q = f - A x = L x - rhs
    const float theta = 0.5f * (8.0f + lmin), delta = 0.5f * (8.0f - lmin);
    const float sigma = theta / delta;
    float rho = 1.0f / sigma;
    float d[CJ_ROWS][2];
    t.lap<true>(x, lv);
#pragma unroll
    for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        r[i][c] = lv[i][c] - r[i][c];
        d[i][c] = r[i][c] / theta;
      }
    for (int s = 0; s < iters; ++s) {
      t.lap<true>(d, lv);
      const float rho1 = 1.0f / (2.0f * sigma - rho);
      const float cd = rho1 * rho, cr = 2.0f * rho1 / delta;
#pragma unroll
      for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          x[i][c] = x[i][c] + d[i][c];
          r[i][c] = r[i][c] + lv[i][c];  // q -= A d
          d[i][c] = __builtin_fmaf(cd, d[i][c], cr * r[i][c]);
        }
      rho = rho1;
    }
  }
#pragma unroll
  for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gx = t.x0 + c, gy = y0 + i;
      if (gx < Wc && gy < Hc) e[gy * Wc + gx] = x[i][c];
    }
}

// Damped Jacobi sweep p <- p + w (J(p) - p), J the plain sweep (same
// operator and summation order).  The plain sweep's iteration matrix has
// eigenvalues (cos kx + cos ky) / 2 in [-1, 1]: the near-checkerboard modes
// decay as slowly as the smoothest ones, and the coarse correction cannot
// reach them (restriction averages them away).  w = 1/2 maps an eigenvalue
// m to (1 + m) / 2, which removes them.
__global__ void __launch_bounds__(256)
    k_damped_sweep(const float* __restrict__ pin, const float* __restrict__ b,
                   const uint8_t* __restrict__ mask, float* __restrict__ pout, int W,
                   int H, float w) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= W || y >= H) return;
  const int pos = y * W + x;
  const int iL = y * W + max(x - 1, 0), iR = y * W + min(x + 1, W - 1);
  const int iB = max(y - 1, 0) * W + x, iT = min(y + 1, H - 1) * W + x;
  const float p = pin[pos];
  const float x1 = mask[iL] ? p : pin[iL];
  const float x2 = mask[iR] ? p : pin[iR];
  const float y1 = mask[iB] ? p : pin[iB];
  const float y2 = mask[iT] ? p : pin[iT];
  const float jac = ((((x1 + x2) + y1) + y2) - b[pos]) * 0.25f;
  pout[pos] = p + w * (jac - p);
}

// The Chebyshev smoothing of k_coarse_smooth<true> spread over 2D tiles:
// a workgroup of CT_WAVES waves holds a 64 x CT_EY tile (lane l column l,
// wave w rows CT_ROWS w .. CT_ROWS w + CT_ROWS - 1), and a launch runs up
// to CB_T iterations on it (plus, at the first launch, the initial
// residual).  A tile overlaps its neighbours by CT_HALO = CB_T + 1 cells on
// every side, whose values go stale one cell per stencil, and stores its
// CT_VX x CT_VY interior: at 128 x 128, 4 x 22 workgroups of 1024 threads,
// 2 cells per lane (the shape and CB_T timed fastest on the 2048^2 learned
// step among 8 to 16 iterations per launch and 4 to 16 waves).  Launch
// boundaries are the only global synchronisation.  Every cell sees the
// operations of the single-workgroup kernel in the same order, so x is
// bitwise the same; wavefronts whose cells have no solid or out-of-grid
// neighbour (the inner tiles away from obstacles) take a select-free
// stencil with the same sums.  Theta and the
// coefficients of the launch's iterations (cf) come from the host, computed
// by the same float32 operations (IEEE single on both sides).
constexpr int CB_T = 12, CT_ROWS = 2, CT_WAVES = 16, CT_HALO = CB_T + 1;
constexpr int CT_EY = CT_ROWS * CT_WAVES, CT_VX = 64 - 2 * CT_HALO, CT_VY = CT_EY - 2 * CT_HALO;
static_assert(CT_VX > 0 && CT_VY > 0, "coarse tile smaller than its halo");

// FACES: the face-weighted operator of k_coarse_faces.  A neighbour across
// a face of weight w enters the sum as the centre c when w = 0 (as a solid
// neighbour does), as itself when w = 1, and as fma(w, v - c, c) between,
// so that the sum minus 4 c is the sum of w (v - c); a cell whose faces
// are all shut or all open sees the operations of the plain kernel.
template <bool FACES>
__global__ void __launch_bounds__(64 * CT_WAVES)
    k_coarse_cheb_tile(const float* __restrict__ xin, const float* __restrict__ rin,
                       const float* __restrict__ din, float* __restrict__ xout,
                       float* __restrict__ rout, float* __restrict__ dout,
                       const float* __restrict__ rhs, const uint8_t* __restrict__ mc,
                       const float* __restrict__ faces, int Wc, int Hc, int s0, int T,
                       float theta, ChebCoef cf) {
  constexpr int R = CT_ROWS;
  const int lane = threadIdx.x & 63;
  // wave-uniform: a scalar register, so branches on it are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gx = (int)blockIdx.x * CT_VX - CT_HALO + lane;
  const int gy0 = (int)blockIdx.y * CT_VY - CT_HALO + wave * R;
  const bool first = s0 == 0;
  // Every load is unconditional, from a clamped in-grid address, issued
  // before the first use: x, r (b at the first launch) and d of the lane's
  // cells, and the mask of its column over the rows gy0 - 1 .. gy0 + R; the
  // neighbours' flags to the left and right come from the adjacent lanes,
  // and a cell outside the grid counts as solid.
  const float* rsrc = first ? rhs : rin;
  const int cx = min(max(gx, 0), Wc - 1);
  float lx[R], lr[R], ld[R];
  uint8_t lm[R + 2];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int pos = min(max(gy0 + i, 0), Hc - 1) * Wc + cx;
    lx[i] = xin[pos];
    lr[i] = rsrc[pos];
    ld[i] = din[pos];
  }
#pragma unroll
  for (int i = 0; i < R + 2; ++i) lm[i] = mc[min(max(gy0 - 1 + i, 0), Hc - 1) * Wc + cx];
  const bool col_in = gx >= 0 && gx < Wc;
  int sf[R + 2];  // solid flags of rows gy0 - 1 + i
#pragma unroll
  for (int i = 0; i < R + 2; ++i) {
    const int gy = gy0 - 1 + i;
    sf[i] = (col_in && gy >= 0 && gy < Hc) ? lm[i] != 0 : 1;
  }
  float x[R], r[R], d[R], lv[R];
  int code[R];  // bit 0: left, 1: right, 2: up (y-1), 3: down (y+1) solid
  // FACES: the weights of the cell's four faces and which are partial
  float wl[FACES ? R : 1], wr[FACES ? R : 1], wu[FACES ? R : 1], wd[FACES ? R : 1];
  int part[FACES ? R : 1];
  float fe[FACES ? R : 1], fs[FACES ? R + 1 : 1];
  if (FACES) {
    const int n = Wc * Hc;
#pragma unroll
    for (int i = 0; i < R; ++i) fe[i] = faces[min(max(gy0 + i, 0), Hc - 1) * Wc + cx];
#pragma unroll
    for (int i = 0; i < R + 1; ++i) fs[i] = faces[n + min(max(gy0 - 1 + i, 0), Hc - 1) * Wc + cx];
  }
  bool any = false;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int gy = gy0 + i;
    const bool in = col_in && gy >= 0 && gy < Hc;
    x[i] = in ? lx[i] : 0.0f;
    r[i] = in ? lr[i] : 0.0f;
    d[i] = (in && !first) ? ld[i] : 0.0f;
    if (FACES) {
      // a cell outside the grid, and a face on the grid's edge, is shut;
      // the west face is the left lane's east face (lane 0 keeps its own:
      // a halo cell)
      wr[i] = (in && gx < Wc - 1) ? fe[i] : 0.0f;
      wd[i] = (in && gy < Hc - 1) ? fs[i + 1] : 0.0f;
      wu[i] = (in && gy > 0) ? fs[i] : 0.0f;
      wl[i] = jt::lane_from_left(wr[i]);
      if (!in) wl[i] = 0.0f;
      code[i] = (wl[i] == 0.0f ? 1 : 0) | (wr[i] == 0.0f ? 2 : 0) | (wu[i] == 0.0f ? 4 : 0) |
                (wd[i] == 0.0f ? 8 : 0);
      part[i] = (wl[i] > 0.0f && wl[i] < 1.0f ? 1 : 0) | (wr[i] > 0.0f && wr[i] < 1.0f ? 2 : 0) |
                (wu[i] > 0.0f && wu[i] < 1.0f ? 4 : 0) | (wd[i] > 0.0f && wd[i] < 1.0f ? 8 : 0);
      any |= (code[i] | part[i]) != 0;
    } else {
      // the lane shifts run with every lane active; lanes 0 and 63 keep
      // their own flag, which only affects halo cells
      const int lnb = jt::ilane_from_left(sf[i + 1]);
      const int rnb = jt::ilane_from_right(sf[i + 1]);
      code[i] = lnb | (rnb << 1) | (sf[i] << 2) | (sf[i + 2] << 3);
      any |= code[i] != 0;
    }
  }
  const bool special = __any(any);  // wave-uniform

  __shared__ float top[2][CT_WAVES][64], bot[2][CT_WAVES][64];
  int buf = 0;
  // out = L v: ((u + l) + r) + d - 4 c, a solid or out-of-grid neighbour
  // replaced by the centre; every thread of the block calls it
  auto lap = [&](const float (&v)[R], float (&out)[R]) {
    top[buf][wave][lane] = v[0];
    bot[buf][wave][lane] = v[R - 1];
    __syncthreads();
    // the rows beyond the tile's first and last are halo: any finite value
    const float up0 = bot[buf][wave > 0 ? wave - 1 : wave][lane];
    const float dn0 = top[buf][wave < CT_WAVES - 1 ? wave + 1 : wave][lane];
    buf ^= 1;
    if (special) {
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const float c = v[i];
        const float fl = jt::lane_from_left(c), fr = jt::lane_from_right(c);
        const float U = i > 0 ? v[i - 1] : up0;
        const float D = i < R - 1 ? v[i + 1] : dn0;
        const int cd = code[i];
        float l = (cd & 1) ? c : fl;
        float rr = (cd & 2) ? c : fr;
        float u = (cd & 4) ? c : U;
        float dd = (cd & 8) ? c : D;
        if (FACES) {
          const int pt = part[i];
          l = (pt & 1) ? __builtin_fmaf(wl[i], fl - c, c) : l;
          rr = (pt & 2) ? __builtin_fmaf(wr[i], fr - c, c) : rr;
          u = (pt & 4) ? __builtin_fmaf(wu[i], U - c, c) : u;
          dd = (pt & 8) ? __builtin_fmaf(wd[i], D - c, c) : dd;
        }
        out[i] = __builtin_fmaf(-4.0f, c, ((u + l) + rr) + dd);
      }
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const float c = v[i];
        const float U = i > 0 ? v[i - 1] : up0;
        const float D = i < R - 1 ? v[i + 1] : dn0;
        // zero-filled shifts at lanes 0 and 63 (halo cells), folded into
        // the additions as DPP operands
        out[i] = __builtin_fmaf(-4.0f, c, ((U + jt::lane_shr0(c)) + jt::lane_shl0(c)) + D);
      }
    }
  };
  if (first) {
    lap(x, lv);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      r[i] = lv[i] - r[i];
      d[i] = r[i] / theta;
    }
  }
  for (int it = 0; it < T; ++it) {
    lap(d, lv);
    const float cd = cf.cd[it], cr = cf.cr[it];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      x[i] = x[i] + d[i];
      r[i] = r[i] + lv[i];
      d[i] = __builtin_fmaf(cd, d[i], cr * r[i]);
    }
  }
  const bool col_out = lane >= CT_HALO && lane < 64 - CT_HALO && gx < Wc;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int ty = wave * R + i, gy = gy0 + i;
    if (col_out && ty >= CT_HALO && ty < CT_EY - CT_HALO && gy < Hc) {
      const int pos = gy * Wc + gx;
      xout[pos] = x[i];
      rout[pos] = r[i];
      dout[pos] = d[i];
    }
  }
}

namespace nx {
int launch_coarse_smooth(hipStream_t st, int Wc, int Hc, float* e, const float* rc,
                         const uint8_t* mc, int iterations, float lambda_min, float* ws,
                         const float* faces) {
  if (iterations == 0) return NX_OK;
  if (faces && !(lambda_min > 0.0f))
    return set_error(NX_ERR_ARG, "the face-weighted coarse operator needs lambda_min > 0");
  // the face-weighted operator always runs tiled (any number of iterations)
  if (lambda_min > 0.0f && (iterations > CB_T || faces)) {
    // tiled: x, r, d ping-pong between the two halves of ws (6 Wc Hc
    // floats); the last launch writes x to e (read only by the first)
    if (!ws) return set_error(NX_ERR_ARG, "coarse smoother workspace missing");
    const size_t n = (size_t)Wc * Hc;
    float* set[2][3] = {{ws, ws + n, ws + 2 * n}, {ws + 3 * n, ws + 4 * n, ws + 5 * n}};
    const dim3 grid((Wc + CT_VX - 1) / CT_VX, (Hc + CT_VY - 1) / CT_VY);
    ChebRecurrence rec(lambda_min);
    int done = 0, cur = 0;
    while (done < iterations) {
      const int T = iterations - done < CB_T ? iterations - done : CB_T;
      const bool last = done + T >= iterations;
      const float* xin = done == 0 ? e : set[cur][0];
      float** out = set[done == 0 ? 0 : 1 - cur];
      ChebCoef cf{};
      rec.next(0, T, cf);
      if (faces)
        k_coarse_cheb_tile<true><<<grid, 64 * CT_WAVES, 0, st>>>(
            xin, set[cur][1], set[cur][2], last ? e : out[0], out[1], out[2], rc, mc, faces,
            Wc, Hc, done, T, rec.theta, cf);
      else
        k_coarse_cheb_tile<false><<<grid, 64 * CT_WAVES, 0, st>>>(
            xin, set[cur][1], set[cur][2], last ? e : out[0], out[1], out[2], rc, mc, nullptr,
            Wc, Hc, done, T, rec.theta, cf);
      if (hipGetLastError() != hipSuccess)
        return set_error(NX_ERR_HIP, "coarse smoother launch failed");
      cur = done == 0 ? 0 : 1 - cur;
      done += T;
    }
  } else if (lambda_min > 0.0f) {
    k_coarse_smooth<true><<<1, 1024, 0, st>>>(e, rc, mc, Wc, Hc, iterations, lambda_min);
  } else {
    k_coarse_smooth<false><<<1, 1024, 0, st>>>(e, rc, mc, Wc, Hc, iterations, 0.0f);
  }
  if (hipGetLastError() != hipSuccess)
    return set_error(NX_ERR_HIP, "coarse smoother launch failed");
  return NX_OK;
}
}  // namespace nx

// Energy line search along a correction d.  With L the 5-point operator of
// the Jacobi sweep (a solid neighbour reads the centre value, an edge
// neighbour is the cell itself), -L is symmetric positive semi-definite on
// the fluid cells and the pressure minimises phi(p) = 1/2 p.(-L)p + b.p.
// Along p + a d, phi is smallest at a = (d.r) / (d.Ld) with r = b - L p, and
// phi(p) - phi(p*) is half the squared energy norm of the error, so the step
// never increases that error.  Per block: FP64 partial sums of d.r, d.Ld.
__global__ void __launch_bounds__(256)
    k_line_terms(const float* __restrict__ p, const float* __restrict__ b,
                 const uint8_t* __restrict__ mask, const float* __restrict__ d,
                 double* __restrict__ red, int W, int H) {
  // column strips as in k_restrict_rows_seg; a solid neighbour is left
  // out (its difference to the centre is 0), a clamped one is the cell itself
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int x = blockIdx.x * 64 + lane, xc = min(x, W - 1);
  const int ex = lane == 0 ? max(x - 1, 0) : min(x + 1, W - 1);
  const int y0 = blockIdx.y * RES_ROWS + wave * RS;
  double v0 = 0.0, v1 = 0.0;
  if (y0 < H) {  // wave-uniform
    const bool last = x + 1 >= W;
    float pv[RS + 2], pe[RS], dv[RS + 2], de[RS], bv[RS];
    int mv[RS + 2], me[RS];
    strip_load(p, pv, pe, y0, xc, ex, W, H);
    strip_load(d, dv, de, y0, xc, ex, W, H);
    strip_load(mask, mv, me, y0, xc, ex, W, H);
#pragma unroll
    for (int i = 0; i < RS; ++i) bv[i] = b[min(y0 + i, H - 1) * W + xc];
#pragma unroll
    for (int i = 0; i < RS; ++i) {
      if (y0 + i >= H) break;  // wave-uniform
      const float pc = pv[i + 1], dc = dv[i + 1];
      const int mcv = mv[i + 1];
      const float pl = nbr_left(pc, pe[i], lane), pr = nbr_right(pc, pe[i], lane, last);
      const float dl = nbr_left(dc, de[i], lane), dr = nbr_right(dc, de[i], lane, last);
      const int ml = inbr_left(mcv, me[i], lane), mr = inbr_right(mcv, me[i], lane, last);
      if (x < W && mcv == 0) {
        const double pcd = (double)pc, dcd = (double)dc;
        const float pn[4] = {pl, pr, pv[i], pv[i + 2]}, dn[4] = {dl, dr, dv[i], dv[i + 2]};
        const int sv[4] = {ml, mr, mv[i], mv[i + 2]};
        double lp = 0.0, ld = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (sv[j]) continue;
          lp += (double)pn[j] - pcd;
          ld += (double)dn[j] - dcd;
        }
        v0 += dcd * ((double)bv[i] - lp);
        v1 += dcd * ld;
      }
    }
  }
  __shared__ double part[4][2];
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
  }
  if (lane == 0) {
    part[wave][0] = v0;
    part[wave][1] = v1;
  }
  __syncthreads();
  if (lane == 0 && wave == 0) {
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    red[2 * bid] = ((part[0][0] + part[1][0]) + part[2][0]) + part[3][0];
    red[2 * bid + 1] = ((part[0][1] + part[1][1]) + part[2][1]) + part[3][1];
  }
}

// line = {a, d.r, d.Ld, 0}; a = 0 when d carries no energy (d.Ld >= 0) or
// either sum is not finite (a NaN or infinite direction)
__global__ void __launch_bounds__(1024)
    k_line_alpha(const double* __restrict__ red, int nblocks, double* __restrict__ line) {
  double v0 = 0.0, v1 = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 1024) {
    v0 += red[2 * i];
    v1 += red[2 * i + 1];
  }
  __shared__ double part[16][2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
  }
  if (lane == 0) {
    part[wave][0] = v0;
    part[wave][1] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double dr = 0.0, dld = 0.0;
    for (int w = 0; w < 16; ++w) {
      dr += part[w][0];
      dld += part[w][1];
    }
    const bool ok = isfinite(dr) && isfinite(dld) && dld < 0.0;
    line[0] = ok ? dr / dld : 0.0;
    line[1] = dr;
    line[2] = dld;
    line[3] = 0.0;
  }
}

// p += a d on fluid cells, a read from the device; a = 0 leaves p untouched
// (d may hold non-finite values then)
__global__ void k_line_apply(float* __restrict__ p, const float* __restrict__ d,
                             const uint8_t* __restrict__ mask,
                             const double* __restrict__ line, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const float a = (float)line[0];
  if (i >= n || mask[i] || a == 0.0f) return;
  p[i] = p[i] + a * d[i];
}

// ---- iterative refinement of the pressure ---------------------------------
// The pressure is held as an unevaluated float32 sum ph + pl.  The outer
// residual r = b0 - A (ph + pl) is evaluated in FP64 (so its accuracy is not
// limited by float32 rounding of a large pressure), stored as the float32
// right-hand side of the correction equation A e = r, and reduced in FP64
// for the relative residual.  The correction e is then solved in float32
// and added to the pair with an error-free two-sum.
__global__ void __launch_bounds__(256)
    k_refine_residual(const float* __restrict__ ph, const float* __restrict__ pl,
                      const float* __restrict__ b0, const uint8_t* __restrict__ mask,
                      float* __restrict__ r_out, float* __restrict__ p1,
                      double* __restrict__ red, int W, int H) {
  // column strips as in k_restrict_rows_seg, with the pair (ph, pl) per cell
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int x = blockIdx.x * 64 + lane, xc = min(x, W - 1);
  const int ex = lane == 0 ? max(x - 1, 0) : min(x + 1, W - 1);
  const int y0 = blockIdx.y * RES_ROWS + wave * RS;
  double v[5] = {0, 0, 0, 0, 0};
  if (y0 < H) {  // wave-uniform
    const bool last = x + 1 >= W;
    float hv[RS + 2], he[RS], lv[RS + 2], le[RS], bv[RS];
    int mv[RS + 2], me[RS];
    strip_load(ph, hv, he, y0, xc, ex, W, H);
    strip_load(pl, lv, le, y0, xc, ex, W, H);
    strip_load(mask, mv, me, y0, xc, ex, W, H);
#pragma unroll
    for (int i = 0; i < RS; ++i) bv[i] = b0[min(y0 + i, H - 1) * W + xc];
#pragma unroll
    for (int i = 0; i < RS; ++i) {
      const int y = y0 + i;
      if (y >= H) break;  // wave-uniform
      const float hc = hv[i + 1], lc = lv[i + 1];
      const int mcv = mv[i + 1];
      const float hl = nbr_left(hc, he[i], lane), hr = nbr_right(hc, he[i], lane, last);
      const float ll = nbr_left(lc, le[i], lane), lr = nbr_right(lc, le[i], lane, last);
      const int ml = inbr_left(mcv, me[i], lane), mr = inbr_right(mcv, me[i], lane, last);
      if (x < W) {
        const int pos = y * W + x;
        const double pc = (double)hc + (double)lc;
        const double x1 = ml ? pc : (double)hl + (double)ll;
        const double x2 = mr ? pc : (double)hr + (double)lr;
        const double y1v = mv[i] ? pc : (double)hv[i] + (double)lv[i];
        const double y2v = mv[i + 2] ? pc : (double)hv[i + 2] + (double)lv[i + 2];
        const double bd = (double)bv[i];
        const double rr = bd - ((((x1 + x2) + y1v) + y2v) - 4.0 * pc);
        const float rf = mcv ? 0.0f : (float)rr;
        r_out[pos] = rf;
        p1[pos] = (0.0f - rf) * 0.25f;
        if (mcv == 0) {
          v[0] += rr;
          v[1] += rr * rr;
          v[2] += bd;
          v[3] += bd * bd;
          v[4] += 1.0;
        }
      }
    }
  }
  __shared__ double part[4][5];
  for (int j = 0; j < 5; ++j)
    for (int off = 32; off > 0; off >>= 1) v[j] += __shfl_xor(v[j], off, 64);
  if (lane == 0)
    for (int j = 0; j < 5; ++j) part[wave][j] = v[j];
  __syncthreads();
  if (lane == 0 && wave == 0) {
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    for (int j = 0; j < 5; ++j)
      red[5 * bid + j] = ((part[0][j] + part[1][j]) + part[2][j]) + part[3][j];
  }
}

// (ph, pl) += e with Knuth's two-sum: s = ph + e, err = (ph - (s - bb)) +
// (e - bb) with bb = s - ph; ph = s, pl += err
__global__ void k_refine_accumulate(float* __restrict__ ph, float* __restrict__ pl,
                                    const float* __restrict__ e, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float a = ph[i], b = e[i];
  const float sum = a + b;
  const float bb = sum - a;
  const float err = (a - (sum - bb)) + (b - bb);
  ph[i] = sum;
  pl[i] = pl[i] + err;
}

__global__ void k_refine_combine(const float* __restrict__ ph, const float* __restrict__ pl,
                                 float* __restrict__ p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = ph[i] + pl[i];
}

// ------------------------------------------------------------- C-ABI
extern "C" {

int nx_smooth(nx_sim* s, int sweeps, float omega, double* out) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->in_step) return fail(NX_ERR_STATE, "nx_smooth outside nx_step_begin/finish");
  if (sweeps < 0 || !(omega >= 0.0f && omega <= 1.0f))
    return fail(NX_ERR_ARG, "sweeps >= 0 and 0 <= omega <= 1 required");
  JacobiCfg c{s->prm.jacobi_k, s->prm.jacobi_rows, s->prm.jacobi_waves, s->prm.jacobi_cols,
              s->prm.jacobi_seg};
  NX_TRY(run_jacobi_tb(s, sweeps, c));
  s->step_sweeps += sweeps;
  if (omega > 0.0f) NX_TRY(nx_damped_sweeps(s, 1, omega));
  NX_HIP(launch_sums(s->p[s->pcur], s->div, s->mask[s->mcur], s->W, s->H, s->red5,
                     s->stream));
  NX_HIP(launch_reduce(s->red5, residual_blocks(s->W, s->H), 5, 0, s->sums, -1.0,
                       nullptr, 0, 0, s->stream, s->host_out_dev, s->line));
  NX_HIP(hipStreamSynchronize(s->stream));
  const double rel = s->host_out[0], alpha = s->host_out[1];
  s->stats.sweeps = s->step_sweeps;
  s->stats.residual = rel;
  s->stats.residual_sweeps = s->step_sweeps;
  if (out) {
    out[0] = rel;
    out[1] = alpha;
  }
  return NX_OK;
}

static int smooth_cheb_impl(nx_sim* s, int iterations, float lambda_min, const int* tile,
                            double* out, bool pending) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (pending && (!s->dcorr || !s->line))
    return fail(NX_ERR_STATE, "no pending line-search step");
  if (!s->in_step) return fail(NX_ERR_STATE, "nx_smooth_chebyshev outside nx_step_begin/finish");
  if (iterations < 1 || !(lambda_min > 0.0f && lambda_min < 8.0f))
    return fail(NX_ERR_ARG, "iterations >= 1 and 0 < lambda_min < 8 required");
  const JacobiCfg c = tile ? JacobiCfg{tile[0], tile[1], tile[2], tile[3]} : s->cheb_cfg;
  const bool stream = c.seg > 0;
  if (stream ? !stream_cfg_supported(c) : !cheb_cfg_supported(c))
    return fail(NX_ERR_ARG, "unsupported Chebyshev tile");
  if (!s->fcheb) NX_HIP(hipMalloc(&s->fcheb, (size_t)4 * s->W * s->H * sizeof(float)));
  // x ping-pongs between the pressure buffers, (r, d) between the halves of
  // fcheb; the first launch spends one pass on the setup, the last one pass
  // on the FP64 residual of the result (records in cheb_red, one per block)
  const size_t n = (size_t)s->W * s->H;
  float* rd[2][2] = {{s->fcheb, s->fcheb + n}, {s->fcheb + 2 * n, s->fcheb + 3 * n}};
  // a launch with fewer passes has a wider interior, so no launch of the
  // run has more workgroups than one with c.k passes
  const size_t red_need =
      (size_t)(stream ? stream_launch_blocks(s->W, s->H, c) : jacobi_launch_blocks(s->W, s->H, c)) * 5;
  if (red_need > s->cheb_red_cap) {
    if (s->cheb_red) NX_HIP(hipFree(s->cheb_red));
    s->cheb_red = nullptr;
    NX_HIP(hipMalloc(&s->cheb_red, red_need * sizeof(double)));
    s->cheb_red_cap = red_need;
  }
  ChebRecurrence rec(lambda_min);
  int done = 0, cur = 0, last_blocks = 0;
  for (bool last = false; !last;) {
    const int setup = done == 0 ? 1 : 0;
    JacobiCfg cc = c;
    last = setup + (iterations - done) + 1 <= c.k;
    cc.k = last ? setup + (iterations - done) + 1 : c.k;
    const int its = cc.k - setup - (last ? 1 : 0);
    ChebCoef cf{};
    rec.next(setup, setup + its, cf);
    if (stream)
      NX_LAUNCH_TIMED(s, NX_TIME_CHEB,
                      launch_cheb_stream(s->stream, cc, s->p[s->pcur], rd[cur][0], rd[cur][1],
                                         s->div, s->mask[s->mcur], s->p[1 - s->pcur],
                                         last ? nullptr : rd[1 - cur][0], rd[1 - cur][1], s->W,
                                         s->H, setup, rec.theta, cf,
                                         last ? s->cheb_red : nullptr,
                                         pending && setup ? s->dcorr : nullptr,
                                         pending && setup ? s->line : nullptr));
    else
      NX_LAUNCH_TIMED(s, NX_TIME_CHEB,
                      launch_cheb_tb(s->stream, cc, s->p[s->pcur], rd[cur][0], rd[cur][1],
                                     s->div, s->mask[s->mcur], s->p[1 - s->pcur],
                                     last ? nullptr : rd[1 - cur][0], rd[1 - cur][1], s->W,
                                     s->H, setup, rec.theta, cf,
                                     last ? s->cheb_red : nullptr,
                                     pending && setup ? s->dcorr : nullptr,
                                     pending && setup ? s->line : nullptr));
    if (last)
      last_blocks = stream ? stream_launch_blocks(s->W, s->H, cc)
                           : jacobi_launch_blocks(s->W, s->H, cc);
    s->pcur ^= 1;
    cur ^= 1;
    done += its;
  }
  s->step_sweeps += iterations;
  // the residual and the line-search step land in pinned host memory,
  // written by the reduction itself
  NX_HIP(launch_reduce(s->cheb_red, last_blocks, 5, 0, s->sums, -1.0, nullptr, 0, 0,
                       s->stream, s->host_out_dev, s->line));
  NX_HIP(hipStreamSynchronize(s->stream));
  const double rel = s->host_out[0], alpha = s->host_out[1];
  s->stats.sweeps = s->step_sweeps;
  s->stats.residual = rel;
  s->stats.residual_sweeps = s->step_sweeps;
  if (out) {
    out[0] = rel;
    out[1] = alpha;
  }
  return NX_OK;
}

int nx_smooth_chebyshev(nx_sim* s, int iterations, float lambda_min, const int* tile,
                        double* out) {
  return smooth_cheb_impl(s, iterations, lambda_min, tile, out, false);
}

int nx_set_smoother_tile(nx_sim* s, const int* cfg, int n) {
  if (!s || !cfg || (n != 4 && n != 5)) return fail(NX_ERR_ARG, "4 or 5 tile values required");
  JacobiCfg c{cfg[0], cfg[1], cfg[2], cfg[3], n == 5 ? cfg[4] : 0};
  if (c.seg < 0 || (c.seg > 0 ? !stream_cfg_supported(c) : !cheb_cfg_supported(c)))
    return fail(NX_ERR_ARG, "unsupported Chebyshev tile");
  s->cheb_cfg = c;
  return NX_OK;
}

int nx_get_smoother_tile(nx_sim* s, int* cfg) {
  if (!s || !cfg) return fail(NX_ERR_ARG, "null argument");
  cfg[0] = s->cheb_cfg.k;
  cfg[1] = s->cheb_cfg.rows;
  cfg[2] = s->cheb_cfg.waves;
  cfg[3] = s->cheb_cfg.cols;
  cfg[4] = s->cheb_cfg.seg;
  return NX_OK;
}

// restriction onto a Wp x Hp coarse grid, Wp >= ceil(W / f), Hp >= ceil(H / f)
// faces != nullptr: block sums and the face weights (2 Wp Hp floats)
static int restrict_impl(nx_sim* s, int factor, int Wp, int Hp, float* coarse_rhs,
                         uint8_t* coarse_mask, float* faces = nullptr) {
  if (!s || !coarse_rhs || !coarse_mask || factor < 1)
    return fail(NX_ERR_ARG, "bad restriction arguments");
  if (!s->in_step) return fail(NX_ERR_STATE, "restriction outside a step");
  const int Wc = (s->W + factor - 1) / factor, Hc = (s->H + factor - 1) / factor;
  if (Wp < Wc || Hp < Hc) return fail(NX_ERR_ARG, "coarse grid smaller than the fine grid / f");
  const size_t need = (size_t)2 * s->H * Wc;
  if (need > s->rrow_cap) {
    if (s->rrow) NX_HIP(hipFree(s->rrow));
    s->rrow = nullptr;
    NX_HIP(hipMalloc(&s->rrow, need * sizeof(float)));
    s->rrow_cap = need;
  }
  float* rsum = s->rrow;
  float* rcnt = s->rrow + (size_t)s->H * Wc;
  if (factor >= 2 && factor <= 64 && (factor & (factor - 1)) == 0)
    k_restrict_rows_seg<<<dim3((s->W + 63) / 64, (s->H + RES_ROWS - 1) / RES_ROWS), dim3(64, 4), 0,
                          s->stream>>>(s->p[s->pcur], s->div, s->mask[s->mcur], rsum, rcnt,
                                       s->W, s->H, factor, Wc);
  else
    k_restrict_rows<<<dim3((Wc + 63) / 64, (s->H + 3) / 4), dim3(64, 4), 0, s->stream>>>(
        s->p[s->pcur], s->div, s->mask[s->mcur], rsum, rcnt, s->W, s->H, factor, Wc);
  NX_HIP(hipGetLastError());
  k_restrict_cols<<<dim3((Wp + 63) / 64, (Hp + 3) / 4), dim3(64, 4), 0, s->stream>>>(
      rsum, rcnt, coarse_rhs, coarse_mask, s->H, factor, Wc, Hc, Wp, Hp, faces ? 1 : 0);
  NX_HIP(hipGetLastError());
  if (faces) {
    k_coarse_faces<<<dim3((Wp + 63) / 64, (Hp + 3) / 4), dim3(64, 4), 0, s->stream>>>(
        s->mask[s->mcur], s->W, s->H, factor, Wp, Hp, faces);
    NX_HIP(hipGetLastError());
  }
  return NX_OK;
}

int nx_restrict_residual(nx_sim* s, int factor, float* coarse_rhs,
                         uint8_t* coarse_mask) {
  if (!s || factor < 1) return fail(NX_ERR_ARG, "bad restriction arguments");
  return restrict_impl(s, factor, (s->W + factor - 1) / factor, (s->H + factor - 1) / factor,
                       coarse_rhs, coarse_mask);
}

int nx_restrict_residual_faces(nx_sim* s, int factor, float* coarse_rhs,
                               uint8_t* coarse_mask, float* faces) {
  if (!s || factor < 1 || !faces) return fail(NX_ERR_ARG, "bad restriction arguments");
  return restrict_impl(s, factor, (s->W + factor - 1) / factor, (s->H + factor - 1) / factor,
                       coarse_rhs, coarse_mask, faces);
}

int nx_prolong_add(nx_sim* s, int factor, const float* coarse,
                   const uint8_t* coarse_mask, float scale) {
  if (!s || !coarse || !coarse_mask || factor < 1)
    return fail(NX_ERR_ARG, "bad prolongation arguments");
  if (!s->in_step) return fail(NX_ERR_STATE, "prolongation outside a step");
  const int Wc = (s->W + factor - 1) / factor, Hc = (s->H + factor - 1) / factor;
  k_prolong<true><<<dim3((s->W + 63) / 64, (s->H + 3) / 4), dim3(64, 4), 0, s->stream>>>(
      s->p[s->pcur], s->mask[s->mcur], coarse, coarse_mask, s->W, s->H, factor, Wc, Hc,
      scale);
  NX_HIP(hipGetLastError());
  return NX_OK;
}

static int coarse_smooth_impl(nx_sim* s, int Wc, int Hc, float* e, const float* rc,
                              const uint8_t* coarse_mask, const float* faces, int iterations,
                              float lambda_min) {
  if (!s || !e || !rc || !coarse_mask) return fail(NX_ERR_ARG, "null argument");
  if (Wc < 1 || Hc < 1 || Wc > CJ_MAX || Hc > CJ_ROWS * CJ_WAVES || iterations < 0)
    return fail(NX_ERR_ARG, "coarse grid must be at most 128 x 128, iterations >= 0");
  if (!(lambda_min < 8.0f)) return fail(NX_ERR_ARG, "lambda_min must be below 8");
  if (faces && !(lambda_min > 0.0f))
    return fail(NX_ERR_ARG, "the face-weighted coarse operator needs lambda_min > 0");
  const size_t n = (size_t)Wc * Hc;
  if (lambda_min > 0.0f && (iterations > CB_T || faces) && 6 * n > s->cws_cap) {
    if (s->cws) NX_HIP(hipFree(s->cws));
    s->cws = nullptr;
    NX_HIP(hipMalloc(&s->cws, 6 * n * sizeof(float)));
    s->cws_cap = 6 * n;
  }
  return nx::launch_coarse_smooth(s->stream, Wc, Hc, e, rc, coarse_mask, iterations,
                                  lambda_min, s->cws, faces);
}

int nx_coarse_smooth(nx_sim* s, int Wc, int Hc, float* e, const float* rc,
                     const uint8_t* coarse_mask, int iterations, float lambda_min) {
  return coarse_smooth_impl(s, Wc, Hc, e, rc, coarse_mask, nullptr, iterations, lambda_min);
}

int nx_coarse_smooth_faces(nx_sim* s, int Wc, int Hc, float* e, const float* rc,
                           const uint8_t* coarse_mask, const float* faces, int iterations,
                           float lambda_min) {
  if (!faces) return fail(NX_ERR_ARG, "null argument");
  return coarse_smooth_impl(s, Wc, Hc, e, rc, coarse_mask, faces, iterations, lambda_min);
}

int nx_damped_sweeps(nx_sim* s, int count, float omega) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->in_step) return fail(NX_ERR_STATE, "damped sweeps outside a step");
  if (count < 0 || !(omega > 0.0f && omega <= 1.0f))
    return fail(NX_ERR_ARG, "count >= 0 and 0 < omega <= 1 required");
  const dim3 grid((s->W + 63) / 64, (s->H + 3) / 4);
  for (int i = 0; i < count; ++i) {
    k_damped_sweep<<<grid, dim3(64, 4), 0, s->stream>>>(
        s->p[s->pcur], s->div, s->mask[s->mcur], s->p[1 - s->pcur], s->W, s->H, omega);
    NX_HIP(hipGetLastError());
    s->pcur ^= 1;
  }
  s->step_sweeps += count;
  return NX_OK;
}

static int ensure_line(nx_sim* s) {
  if (s->dcorr) return NX_OK;
  NX_HIP(hipMalloc(&s->dcorr, (size_t)s->W * s->H * sizeof(float)));
  NX_HIP(hipMalloc(&s->line, 4 * sizeof(double)));
  return NX_OK;
}

static int line_search(nx_sim* s, const float* d, double* out, bool apply = true) {
  const size_t n = (size_t)s->W * s->H;
  const dim3 grid((s->W + 63) / 64, (s->H + RES_ROWS - 1) / RES_ROWS);
  // red5 holds 5 doubles per block of this grid; the line terms use 2
  k_line_terms<<<grid, dim3(64, 4), 0, s->stream>>>(s->p[s->pcur], s->div, s->mask[s->mcur],
                                                    d, s->red5, s->W, s->H);
  NX_HIP(hipGetLastError());
  k_line_alpha<<<1, 1024, 0, s->stream>>>(s->red5, residual_blocks(s->W, s->H), s->line);
  NX_HIP(hipGetLastError());
  if (apply) {
    k_line_apply<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(
        s->p[s->pcur], d, s->mask[s->mcur], s->line, (int)n);
    NX_HIP(hipGetLastError());
  }
  if (out) {
    NX_HIP(hipMemcpyAsync(out, s->line, 3 * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
    NX_HIP(hipStreamSynchronize(s->stream));
  }
  return NX_OK;
}

int nx_line_search(nx_sim* s, const float* d, double* out) {
  if (!s || !d) return fail(NX_ERR_ARG, "null argument");
  if (!s->in_step) return fail(NX_ERR_STATE, "line search outside a step");
  NX_TRY(ensure_line(s));
  return line_search(s, d, out);
}

// interpolation from a Wp x Hp coarse grid (padded cells must be solid) and
// the energy line search
static int prolong_ls_impl(nx_sim* s, int factor, int Wp, int Hp, const float* coarse,
                           const uint8_t* coarse_mask, double* out, bool apply = true) {
  if (!s || !coarse || !coarse_mask || factor < 1)
    return fail(NX_ERR_ARG, "bad prolongation arguments");
  if (!s->in_step) return fail(NX_ERR_STATE, "prolongation outside a step");
  const int Wc = (s->W + factor - 1) / factor, Hc = (s->H + factor - 1) / factor;
  if (Wp < Wc || Hp < Hc) return fail(NX_ERR_ARG, "coarse grid smaller than the fine grid / f");
  NX_TRY(ensure_line(s));
  const dim3 grid((s->W + 63) / 64, (s->H + 3) / 4);
  k_prolong<false><<<grid, dim3(64, 4), 0, s->stream>>>(
      s->dcorr, s->mask[s->mcur], coarse, coarse_mask, s->W, s->H, factor, Wp, Hp, 1.0f);
  NX_HIP(hipGetLastError());
  return line_search(s, s->dcorr, out, apply);
}

int nx_prolong_line_search(nx_sim* s, int factor, const float* coarse,
                           const uint8_t* coarse_mask, double* out) {
  if (!s || factor < 1) return fail(NX_ERR_ARG, "bad prolongation arguments");
  return prolong_ls_impl(s, factor, (s->W + factor - 1) / factor,
                         (s->H + factor - 1) / factor, coarse, coarse_mask, out);
}

int nx_refine_begin(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->in_step) return fail(NX_ERR_STATE, "refinement outside a step");
  const size_t n = (size_t)s->W * s->H;
  if (!s->ph) {
    NX_HIP(hipMalloc(&s->ph, n * sizeof(float)));
    NX_HIP(hipMalloc(&s->pl, n * sizeof(float)));
    NX_HIP(hipMalloc(&s->b0, n * sizeof(float)));
  }
  NX_HIP(hipMemcpyAsync(s->b0, s->div, n * sizeof(float), hipMemcpyDeviceToDevice,
                        s->stream));
  NX_HIP(launch_fill(s->ph, 0.0f, n, s->stream));
  NX_HIP(launch_fill(s->pl, 0.0f, n, s->stream));
  s->refining = true;
  return NX_OK;
}

int nx_refine_residual(nx_sim* s, double* rel) {
  if (!s || !rel) return fail(NX_ERR_ARG, "null argument");
  if (!s->refining) return fail(NX_ERR_STATE, "nx_refine_begin first");
  const dim3 grid((s->W + 63) / 64, (s->H + RES_ROWS - 1) / RES_ROWS);
  k_refine_residual<<<grid, dim3(64, 4), 0, s->stream>>>(
      s->ph, s->pl, s->b0, s->mask[s->mcur], s->div, s->p[0], s->red5, s->W, s->H);
  NX_HIP(hipGetLastError());
  NX_HIP(launch_reduce(s->red5, residual_blocks(s->W, s->H), 5, 0, s->sums, -1.0,
                       nullptr, 0, 0, s->stream, s->host_out_dev));
  NX_HIP(hipStreamSynchronize(s->stream));
  s->pcur = 0;
  *rel = s->host_out[0];
  return NX_OK;
}

int nx_refine_accumulate(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->refining) return fail(NX_ERR_STATE, "nx_refine_begin first");
  const int n = s->W * s->H;
  k_refine_accumulate<<<(n + 255) / 256, 256, 0, s->stream>>>(s->ph, s->pl,
                                                              s->p[s->pcur], n);
  NX_HIP(hipGetLastError());
  return NX_OK;
}

int nx_refine_end(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->refining) return fail(NX_ERR_STATE, "nx_refine_begin first");
  const int n = s->W * s->H;
  // the projection reads the pressure ph + pl; the right-hand side slot
  // gets the original (centred) divergence back
  k_refine_combine<<<(n + 255) / 256, 256, 0, s->stream>>>(s->ph, s->pl, s->p[0], n);
  NX_HIP(hipGetLastError());
  NX_HIP(hipMemcpyAsync(s->div, s->b0, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice,
                        s->stream));
  s->pcur = 0;
  s->refining = false;
  return NX_OK;
}

}  // extern "C"

namespace nx {
int restrict_residual_padded(nx_sim* s, int factor, int Wp, int Hp, float* rc, uint8_t* mc,
                             float* faces) {
  return restrict_impl(s, factor, Wp, Hp, rc, mc, faces);
}

int block_sum_restrict(hipStream_t st, const float* rm, const uint8_t* mm, int Wm, int Hm, int f,
                       float* rc, uint8_t* mc, int Wc, int Hc) {
  k_block_sum<<<dim3((Wc + 63) / 64, (Hc + 3) / 4), dim3(64, 4), 0, st>>>(rm, mm, Wm, Hm, f, rc,
                                                                         mc, Wc, Hc);
  return hipGetLastError() == hipSuccess ? NX_OK : set_error(NX_ERR_HIP, "block sum failed");
}

int coarse_faces_padded(nx_sim* s, int f, int Wp, int Hp, float* faces) {
  k_coarse_faces<<<dim3((Wp + 63) / 64, (Hp + 3) / 4), dim3(64, 4), 0, s->stream>>>(
      s->mask[s->mcur], s->W, s->H, f, Wp, Hp, faces);
  return hipGetLastError() == hipSuccess ? NX_OK : set_error(NX_ERR_HIP, "faces failed");
}

int prolong_to(hipStream_t st, float* e, const uint8_t* m, int W, int H, int f, const float* ec,
               const uint8_t* mc, int Wc, int Hc) {
  k_prolong<false><<<dim3((W + 63) / 64, (H + 3) / 4), dim3(64, 4), 0, st>>>(
      e, m, ec, mc, W, H, f, Wc, Hc, 1.0f);
  return hipGetLastError() == hipSuccess ? NX_OK : set_error(NX_ERR_HIP, "prolongation failed");
}
int prolong_line_search_padded(nx_sim* s, int factor, int Wp, int Hp, const float* e,
                               const uint8_t* mc, double* out, bool apply) {
  return prolong_ls_impl(s, factor, Wp, Hp, e, mc, out, apply);
}
int smooth_chebyshev_pending(nx_sim* s, int iterations, float lambda_min, const int* tile,
                             double* out) {
  return smooth_cheb_impl(s, iterations, lambda_min, tile, out, true);
}
}  // namespace nx
