// Temporal-blocked Chebyshev smoothing of the fine pressure system.
//
// Chebyshev iteration for -L x = -b on the eigenvalue interval [lmin, 8] of
// -L (three-term recurrence of Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Algorithm 12.1), with L the 5-point operator of the
// Jacobi sweep (shader.Poisson.comp:26-37: a solid or out-of-grid neighbour
// reads the centre value):
//   setup:      r = L x - b,  d = r / theta
//   iteration:  x += d,  r += L d,  d = (rho' rho) d + (2 rho' / delta) r
//               (one fused multiply-add: fma(cd, d, cr r))
// with theta = (8 + lmin) / 2, delta = (8 - lmin) / 2, sigma = theta /
// delta, rho_0 = 1 / sigma, rho' = 1 / (2 sigma - rho) (the scalar
// recurrence runs on the host; a launch takes its passes' coefficients).
// -L is symmetric positive semi-definite with spectrum in [0, 8], so the
// error polynomial is at most 1 in magnitude on all of [0, 8] and at most
// 1 / T_n(sigma) on [lmin, 8]: n iterations damp every mode above lmin,
// which n Jacobi sweeps (factor (1 - lambda / 4)^n) leave nearly untouched
// near lmin.
//
// Only d needs neighbours, so a launch carries (x, r, d) of a register tile
// through k stencil passes exactly as the Jacobi tile carries p: lane l of
// wave w owns COLS columns and ROWS rows, horizontal neighbours come by DPP
// lane shifts, the rows next to a wave's strip through a double-buffered
// LDS exchange, and the (64 COLS - 2k) x (ROWS WAVES - 2k) interior whose
// k-pass dependency cone stayed inside the tile is written back.  The first
// launch of a run spends its first pass on the setup, the last one (RES)
// its last pass on the FP64 residual of the result.  Every pass evaluates
// each cell's expressions in a fixed order: L v = (((u + l) + r) + d) - 4 v
// with u, d the cells above and below and l, r the lane-shifted ones, so
// that each shift is the DPP operand of an addition (no separate move); the
// -4 v is one fused multiply-add, which rounds like the subtraction because
// 4 v is exact; and the d update is the one fused multiply-add written
// above.  The result is bitwise independent of the tile shape and of k.
#include <type_traits>

#include "../../include/natrix.h"
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"

namespace nx {

namespace {

using jt::lane_from_left;
using jt::lane_from_right;
using jt::ilane_from_left;
using jt::ilane_from_right;

// RES: the launch's last pass evaluates the exact FP64 residual
// r = b - L x of the final iterate on the written interior (the expression
// of the residual sums, r = b - ((((x1 + x2) + y1) + y2) - 4 x) in double)
// and writes red[5 blk + {sum r, sum r^2, sum b, sum b^2, fluid cells}].
template <int ROWS, int WAVES, int COLS, bool RES>
__global__ void __launch_bounds__(WAVES * 64)
    __attribute__((amdgpu_waves_per_eu(jt::cheb_waves_per_eu<ROWS * COLS, RES>())))
    k_cheb_tb(const float* __restrict__ xin, const float* __restrict__ rin,
              const float* __restrict__ din, const float* __restrict__ b,
              const uint8_t* __restrict__ mask, float* __restrict__ xout,
              float* __restrict__ rout, float* __restrict__ dout, int W, int H, int k,
              int first, float theta, ChebCoef cf, double* __restrict__ red,
              const float* __restrict__ dapp, const double* __restrict__ line) {
  constexpr int EX = 64 * COLS, EY = ROWS * WAVES;
  const int lane = threadIdx.x & 63;
  // wave-uniform: a scalar register, so branches on it are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int VX = EX - 2 * k, VY = EY - 2 * k;
  // 1D grid of gx * gy tiles in XCD order (xcd_tile); the partial-sum
  // record is the tile's row-major index
  const int gx = (W + VX - 1) / VX, gy = (H + VY - 1) / VY;
  const int tile = xcd_tile<XCD_RUNS>((int)blockIdx.x, gx * gy);
  const int gx0 = (tile % gx) * VX - k + COLS * lane;
  const int gy0 = (tile / gx) * VY - k + wave * ROWS;
  const int tcol0 = COLS * lane, trow0 = wave * ROWS;

  // float2 accesses when every lane's column pair is aligned and wholly
  // inside or outside the grid and the written interior
  const bool vec = COLS == 2 && (W % 2 == 0) && (k % 2 == 0);
  float x[ROWS][COLS], r[ROWS][COLS], d[ROWS][COLS];
  float bb[RES ? ROWS : 1][RES ? COLS : 1];  // b for the residual pass
  int sol[ROWS][COLS];
  // Every load is unconditional, from a clamped in-grid address, and all of
  // a lane's loads are issued before the first is used; cells outside the
  // grid are replaced afterwards.  The first launch holds b in r until the
  // setup pass (its d input is never read: d is set by the setup).
  const float* rsrc = first ? b : rin;
  const bool apply = dapp != nullptr && (float)line[0] != 0.0f;
  const float a_line = apply ? (float)line[0] : 0.0f;
  if (vec) {
    const int cx = min(max(gx0, 0), W - 2);  // even, like gx0
    const bool col_in = gx0 >= 0 && gx0 < W;
    float2 lx[ROWS], lr[ROWS], ld[ROWS], lb[RES ? ROWS : 1];
    uint16_t lm[ROWS];
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int idx = min(max(gy0 + i, 0), H - 1) * W + cx;
      lx[i] = *reinterpret_cast<const float2*>(xin + idx);
      lr[i] = *reinterpret_cast<const float2*>(rsrc + idx);
      ld[i] = *reinterpret_cast<const float2*>(din + idx);
      lm[i] = *reinterpret_cast<const uint16_t*>(mask + idx);
      if (RES) lb[RES ? i : 0] = *reinterpret_cast<const float2*>(b + idx);
    }
    float2 la[ROWS];
    if (apply) {
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
        la[i] = *reinterpret_cast<const float2*>(dapp + min(max(gy0 + i, 0), H - 1) * W + cx);
    }
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const bool in = col_in && gy0 + i >= 0 && gy0 + i < H;
      x[i][0] = in ? lx[i].x : 0.0f;
      x[i][COLS - 1] = in ? lx[i].y : 0.0f;
      r[i][0] = in ? lr[i].x : 0.0f;
      r[i][COLS - 1] = in ? lr[i].y : 0.0f;
      d[i][0] = (in && !first) ? ld[i].x : 0.0f;
      d[i][COLS - 1] = (in && !first) ? ld[i].y : 0.0f;
      sol[i][0] = in ? (lm[i] & 0xFF) != 0 : 1;
      sol[i][COLS - 1] = in ? (lm[i] >> 8) != 0 : 1;
      if (RES) {
        bb[RES ? i : 0][0] = in ? lb[RES ? i : 0].x : 0.0f;
        bb[RES ? i : 0][RES ? COLS - 1 : 0] = in ? lb[RES ? i : 0].y : 0.0f;
      }
      // a pending line-search step x += a d on the fluid cells, as the
      // apply of the line search would have stored it (first launch only)
      if (apply) {
        if (in && !sol[i][0]) x[i][0] = x[i][0] + a_line * la[i].x;
        if (in && !sol[i][COLS - 1]) x[i][COLS - 1] = x[i][COLS - 1] + a_line * la[i].y;
      }
    }
  } else {
    float lx[ROWS][COLS], lr[ROWS][COLS], ld[ROWS][COLS], lb[RES ? ROWS : 1][RES ? COLS : 1];
    uint8_t lm[ROWS][COLS];
#pragma unroll
    for (int i = 0; i < ROWS; ++i)
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int idx = min(max(gy0 + i, 0), H - 1) * W + min(max(gx0 + c, 0), W - 1);
        lx[i][c] = xin[idx];
        lr[i][c] = rsrc[idx];
        ld[i][c] = din[idx];
        lm[i][c] = mask[idx];
        if (RES) lb[RES ? i : 0][RES ? c : 0] = b[idx];
      }
    float la[ROWS][COLS];
    if (apply) {
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c)
          la[i][c] = dapp[min(max(gy0 + i, 0), H - 1) * W + min(max(gx0 + c, 0), W - 1)];
    }
#pragma unroll
    for (int i = 0; i < ROWS; ++i)
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gy = gy0 + i, gx = gx0 + c;
        const bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
        x[i][c] = in ? lx[i][c] : 0.0f;
        r[i][c] = in ? lr[i][c] : 0.0f;
        d[i][c] = (in && !first) ? ld[i][c] : 0.0f;
        sol[i][c] = in ? (lm[i][c] != 0) : 1;
        if (RES) bb[RES ? i : 0][RES ? c : 0] = in ? lb[RES ? i : 0][RES ? c : 0] : 0.0f;
        if (apply && in && !sol[i][c]) x[i][c] = x[i][c] + a_line * la[i][c];
      }
  }

  __shared__ float xch[2][2][WAVES][EX];
  __shared__ int sxch[2][WAVES][EX];
#pragma unroll
  for (int c = 0; c < COLS; ++c) {
    sxch[0][wave][tcol0 + c] = sol[0][c];
    sxch[1][wave][tcol0 + c] = sol[ROWS - 1][c];
  }
  __syncthreads();
  // bit0 L, bit1 R, bit2 B (row above in memory, y - 1), bit3 T (y + 1):
  // that neighbour is solid or outside the grid and reads the centre
  int cd[ROWS][COLS];
  bool any_nbr = false;
#pragma unroll
  for (int i = 0; i < ROWS; ++i) {
    const int sl = ilane_from_left(sol[i][COLS - 1]);
    const int sr = ilane_from_right(sol[i][0]);
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const int L = c > 0 ? sol[i][c - 1] : sl;
      const int R = c < COLS - 1 ? sol[i][c + 1] : sr;
      // the rows beyond the tile's first and last count as fluid: their
      // cells are outside the valid cone, and the waves holding them then
      // take the fast path too
      const int B = i > 0 ? sol[i - 1][c] : (wave > 0 ? sxch[1][wave - 1][tcol0 + c] : 0);
      const int T = i < ROWS - 1 ? sol[i + 1][c]
                                 : (wave < WAVES - 1 ? sxch[0][wave + 1][tcol0 + c] : 0);
      cd[i][c] = L | (R << 1) | (B << 2) | (T << 3);
      any_nbr |= cd[i][c] != 0;
    }
  }
  // No solid or out-of-grid neighbour anywhere in the wave's cells: the
  // passes run a copy of the stencil with no neighbour replacement.  The
  // flag is wave-uniform, so the choice is one scalar branch per launch.
  const bool special = __any(any_nbr);
  // the rows above and below a wave's strip come from its neighbours in the
  // block; the first and last waves read their own published rows instead
  // (tile rows outside every pass's valid cone, so any finite value serves),
  // which keeps the LDS reads unconditional
  const int wave_up = wave > 0 ? wave - 1 : 0;
  const int wave_dn = wave < WAVES - 1 ? wave + 1 : WAVES - 1;

  int s = 0;
  // passes with (SP = true) or without neighbour replacement
  auto run = [&](auto spc) {
    constexpr bool SP = decltype(spc)::value;
    // f(i, c, centre, x1, x2, y1, y2) for every tile cell with the
    // neighbours of v (garbage on the tile's edge cells, which lie outside
    // the valid cone: lane 0's left and lane 63's right neighbour read 0)
    auto stencil = [&](const float (&v)[ROWS][COLS], int buf, auto&& f) {
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        xch[buf][0][wave][tcol0 + c] = v[0][c];
        xch[buf][1][wave][tcol0 + c] = v[ROWS - 1][c];
      }
      __syncthreads();
      float prev[COLS], above[COLS];
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        prev[c] = xch[buf][1][wave_up][tcol0 + c];
        above[c] = xch[buf][0][wave_dn][tcol0 + c];
      }
#pragma unroll
      for (int i = 0; i < ROWS; ++i) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const float vc = v[i][c];
          float x1 = c > 0 ? v[i][c - 1] : jt::lane_shr0(v[i][COLS - 1]);
          float x2 = c < COLS - 1 ? v[i][c + 1] : jt::lane_shl0(v[i][0]);
          float y1 = i > 0 ? v[i - 1][c] : prev[c];
          float y2 = i < ROWS - 1 ? v[i + 1][c] : above[c];
          if (SP) {
            const int code = cd[i][c];
            x1 = (code & 1) ? vc : x1;
            x2 = (code & 2) ? vc : x2;
            y1 = (code & 4) ? vc : y1;
            y2 = (code & 8) ? vc : y2;
          }
          f(i, c, vc, x1, x2, y1, y2);
        }
      }
    };
    // lv = L v = ((((y1 + x1) + x2) + y2) - 4 v): 4 v is exact, so the
    // fused multiply-add rounds once, like the subtraction
    float lv[ROWS][COLS];
    auto lap = [&](const float (&v)[ROWS][COLS], int buf) {
      stencil(v, buf, [&](int i, int c, float vc, float x1, float x2, float y1, float y2) {
        lv[i][c] = __builtin_fmaf(-4.0f, vc, ((y1 + x1) + x2) + y2);
      });
    };

    if (first) {
      lap(x, 0);
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          r[i][c] = lv[i][c] - r[i][c];
          d[i][c] = r[i][c] / theta;
        }
      s = 1;
    }
    const int iter_end = RES ? k - 1 : k;
    auto iterate = [&](int it) {
      lap(d, it & 1);
      const float cdd = cf.cd[it], crr = cf.cr[it];
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          x[i][c] = x[i][c] + d[i][c];
          r[i][c] = r[i][c] + lv[i][c];
          d[i][c] = __builtin_fmaf(cdd, d[i][c], crr * r[i][c]);
        }
    };
    // two iterations per trip: the fused multiply-add leaves the new d in
    // the product's register, and with two copies of the body the register
    // allocator alternates the roles instead of copying d back every pass
    for (; s + 1 < iter_end; s += 2) {
      iterate(s);
      iterate(s + 1);
    }
    if (s < iter_end) iterate(s++);
  };
  if (special)
    run(std::true_type());
  else
    run(std::false_type());

  if (RES) {
    // the residual pass, with neighbour replacement wherever it applies
    auto stencil = [&](const float (&v)[ROWS][COLS], int buf, auto&& f) {
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        xch[buf][0][wave][tcol0 + c] = v[0][c];
        xch[buf][1][wave][tcol0 + c] = v[ROWS - 1][c];
      }
      __syncthreads();
      float prev[COLS], above[COLS];
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        prev[c] = xch[buf][1][wave_up][tcol0 + c];
        above[c] = xch[buf][0][wave_dn][tcol0 + c];
      }
#pragma unroll
      for (int i = 0; i < ROWS; ++i) {
        const float lft = lane_from_left(v[i][COLS - 1]);
        const float rgt = lane_from_right(v[i][0]);
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const float vc = v[i][c];
          const int code = cd[i][c];
          const float x1 = (code & 1) ? vc : (c > 0 ? v[i][c - 1] : lft);
          const float x2 = (code & 2) ? vc : (c < COLS - 1 ? v[i][c + 1] : rgt);
          const float y1 = (code & 4) ? vc : (i > 0 ? v[i - 1][c] : prev[c]);
          const float y2 = (code & 8) ? vc : (i < ROWS - 1 ? v[i + 1][c] : above[c]);
          f(i, c, vc, x1, x2, y1, y2);
        }
      }
    };
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    stencil(x, s & 1, [&](int i, int c, float vc, float x1, float x2, float y1, float y2) {
      const int trow = trow0 + i, tcol = tcol0 + c, gy = gy0 + i, gx = gx0 + c;
      if (trow < k || trow >= EY - k || tcol < k || tcol >= EX - k || gy < 0 || gy >= H ||
          gx < 0 || gx >= W || sol[i][c])
        return;
      const double bd = (double)bb[RES ? i : 0][RES ? c : 0];
      const double rr = bd - (((((double)x1 + (double)x2) + (double)y1) + (double)y2) -
                              4.0 * (double)vc);
      acc[0] += rr;
      acc[1] += rr * rr;
      acc[2] += bd;
      acc[3] += bd * bd;
      acc[4] += 1.0;
    });
    __shared__ double part[WAVES][5];
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] = jt::wave_sum(acc[j]);
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < 5; ++j) part[wave][j] = acc[j];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int j = 0; j < 5; ++j) {
        double t = 0.0;
        for (int w = 0; w < WAVES; ++w) t += part[w][j];
        red[5 * tile + j] = t;
      }
    }
  }

#pragma unroll
  for (int i = 0; i < ROWS; ++i) {
    const int trow = trow0 + i, gy = gy0 + i;
    if (trow < k || trow >= EY - k || gy < 0 || gy >= H) continue;
    if (vec) {
      if (tcol0 < k || tcol0 >= EX - k || gx0 < 0 || gx0 >= W) continue;
      const int idx = gy * W + gx0;
      *reinterpret_cast<float2*>(xout + idx) = make_float2(x[i][0], x[i][COLS - 1]);
      if (rout) {
        *reinterpret_cast<float2*>(rout + idx) = make_float2(r[i][0], r[i][COLS - 1]);
        *reinterpret_cast<float2*>(dout + idx) = make_float2(d[i][0], d[i][COLS - 1]);
      }
      continue;
    }
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const int tcol = tcol0 + c, gx = gx0 + c;
      if (tcol < k || tcol >= EX - k || gx < 0 || gx >= W) continue;
      const int idx = gy * W + gx;
      xout[idx] = x[i][c];
      if (rout) {
        rout[idx] = r[i][c];
        dout[idx] = d[i][c];
      }
    }
  }
}

template <int ROWS, int WAVES, int COLS>
hipError_t launch_cheb(const float* xin, const float* rin, const float* din, const float* b,
                       const uint8_t* mask, float* xout, float* rout, float* dout, int W,
                       int H, int k, int first, float theta, const ChebCoef& cf, double* red,
                       const float* dapp, const double* line, hipStream_t st) {
  constexpr int EX = 64 * COLS, EY = ROWS * WAVES;
  const dim3 grid(((W + EX - 2 * k - 1) / (EX - 2 * k)) * ((H + EY - 2 * k - 1) / (EY - 2 * k)));
  if (red)
    k_cheb_tb<ROWS, WAVES, COLS, true><<<grid, WAVES * 64, 0, st>>>(
        xin, rin, din, b, mask, xout, rout, dout, W, H, k, first, theta, cf, red, dapp, line);
  else
    k_cheb_tb<ROWS, WAVES, COLS, false><<<grid, WAVES * 64, 0, st>>>(
        xin, rin, din, b, mask, xout, rout, dout, W, H, k, first, theta, cf, nullptr, dapp,
        line);
  return hipGetLastError();
}

}  // namespace

// Tile shapes (rows per wave, waves per block, columns per lane).
#define NX_CHEB_SHAPES(X) \
  X(2, 16, 2)             \
  X(4, 8, 2)              \
  X(4, 16, 2)             \
  X(4, 16, 1)             \
  X(8, 8, 1)              \
  X(8, 16, 1)

bool cheb_cfg_supported(JacobiCfg c) {
  bool shape = false;
#define NX_CHEB_CHECK(R, WV, CL) shape |= (c.rows == R && c.waves == WV && c.cols == CL);
  NX_CHEB_SHAPES(NX_CHEB_CHECK)
#undef NX_CHEB_CHECK
  return shape && c.k >= 1 && c.k <= CHEB_MAX_PASSES && 64 * c.cols - 2 * c.k >= 8 &&
         c.rows * c.waves - 2 * c.k >= 8;
}

hipError_t launch_cheb_tb(hipStream_t st, JacobiCfg c, const float* xin, const float* rin,
                          const float* din, const float* b, const uint8_t* mask, float* xout,
                          float* rout, float* dout, int W, int H, int setup, float theta,
                          const ChebCoef& cf, double* red, const float* dapp,
                          const double* line) {
  if (!cheb_cfg_supported(c)) return hipErrorInvalidValue;
#define NX_CHEB_LAUNCH(R, WV, CL)                                                     \
  if (c.rows == R && c.waves == WV && c.cols == CL)                                  \
    return launch_cheb<R, WV, CL>(xin, rin, din, b, mask, xout, rout, dout, W, H, c.k, \
                                  setup, theta, cf, red, dapp, line, st);
  NX_CHEB_SHAPES(NX_CHEB_LAUNCH)
#undef NX_CHEB_LAUNCH
  return hipErrorInvalidValue;
}

ChebRecurrence::ChebRecurrence(float lambda_min) {
  theta = 0.5f * (8.0f + lambda_min);
  delta = 0.5f * (8.0f - lambda_min);
  sigma = theta / delta;
  rho = 1.0f / sigma;
}

void ChebRecurrence::next(int first, int passes, ChebCoef& cf) {
  for (int i = first; i < passes; ++i) {
    const float rho1 = 1.0f / (2.0f * sigma - rho);
    cf.cd[i] = rho1 * rho;
    cf.cr[i] = 2.0f * rho1 / delta;
    rho = rho1;
  }
}

}  // namespace nx
