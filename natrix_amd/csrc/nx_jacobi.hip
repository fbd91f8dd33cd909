// Temporal-blocked Jacobi pressure sweeps and FP64 residual reductions.
//
// One launch performs k consecutive sweeps of shader.Poisson.comp:24-37 on a
// tile held in registers.  A block is WAVES wavefronts stacked vertically;
// each lane owns one x column and ROWS consecutive rows of it, so:
//   * vertical neighbours inside a lane's strip are registers,
//   * horizontal neighbours are the adjacent lanes' registers, read with the
//     DPP wave_shr:1 / wave_shl:1 lane shifts (no LDS traffic),
//   * the rows just above/below a wave's strip come from the neighbouring
//     wave through a small double-buffered LDS exchange (one barrier/sweep).
// The block reads a (64) x (ROWS*WAVES) tile once, sweeps k times, and writes
// back the (64-2k) x (ROWS*WAVES-2k) interior whose k-step dependency cone
// stayed inside the tile.  Each sweep evaluates exactly the reference's
// expression for every cell, so k sweeps per launch are bitwise equal to k
// single sweeps.
//
// Out-of-grid tile cells and the reference's clamp-to-edge neighbours
// (common.sh:14-17) are both expressed as "use the centre value", which is
// also what a solid neighbour does (Poisson.comp:32-35): a 4-bit code per
// cell says which neighbours are replaced (bit 4: the cell is not fluid).  Wavefronts whose codes are all
// zero (the bulk of the grid) take a select-free path.
#include "nx_kernels.h"

namespace nx {

namespace {

__device__ __forceinline__ float lane_from_left(float v) {
  // lane i receives lane i-1 (wave_shr:1); lane 0 keeps its own value
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138,
                                  0xf, 0xf, false));
}

__device__ __forceinline__ float lane_from_right(float v) {
  // lane i receives lane i+1 (wave_shl:1); lane 63 keeps its own value
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130,
                                  0xf, 0xf, false));
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int ilane_from_left(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}

__device__ __forceinline__ int ilane_from_right(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false);
}

// ROWS x COLS cells per lane: lane `l` of wave `w` owns tile columns
// [COLS*l, COLS*l + COLS) and tile rows [ROWS*w, ROWS*w + ROWS).
template <int ROWS, int WAVES, int COLS>
__global__ void __launch_bounds__(WAVES * 64)
    k_jacobi_tb(const float* __restrict__ pin, const float* __restrict__ b,
                const uint8_t* __restrict__ mask, float* __restrict__ pout,
                int W, int H, int k, double* __restrict__ red,
                const int* __restrict__ done) {
  if (done != nullptr && *done != 0) return;  // block-uniform early exit
  constexpr int EX = 64 * COLS, EY = ROWS * WAVES;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int VX = EX - 2 * k, VY = EY - 2 * k;
  const int gx0 = (int)blockIdx.x * VX - k + COLS * lane;
  const int gy0 = (int)blockIdx.y * VY - k + wave * ROWS;
  // float2 loads/stores when every lane's column pair is aligned and either
  // wholly inside or wholly outside the grid and the valid band
  const bool vec = COLS == 2 && (W % 2 == 0) && (k % 2 == 0);

  float p[ROWS][COLS], bb[ROWS][COLS];
  int sol[ROWS][COLS];  // cell is solid or outside the grid
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int gy = gy0 + r;
    const bool row_in = gy >= 0 && gy < H;
    if (vec) {
      const bool in = row_in && gx0 >= 0 && gx0 < W;
      float2 pv = make_float2(0.0f, 0.0f), bv = pv;
      int m0 = 1, m1 = 1;
      if (in) {
        const int idx = gy * W + gx0;
        pv = *reinterpret_cast<const float2*>(pin + idx);
        bv = *reinterpret_cast<const float2*>(b + idx);
        const uint16_t mm = *reinterpret_cast<const uint16_t*>(mask + idx);
        m0 = (mm & 0xFF) != 0;
        m1 = (mm >> 8) != 0;
      }
      p[r][0] = pv.x;
      p[r][COLS - 1] = pv.y;
      bb[r][0] = bv.x;
      bb[r][COLS - 1] = bv.y;
      sol[r][0] = m0;
      sol[r][COLS - 1] = m1;
    } else {
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gx = gx0 + c;
        p[r][c] = 0.0f;
        bb[r][c] = 0.0f;
        sol[r][c] = 1;
        if (row_in && gx >= 0 && gx < W) {
          const int idx = gy * W + gx;
          p[r][c] = pin[idx];
          bb[r][c] = b[idx];
          sol[r][c] = mask[idx] != 0;
        }
      }
    }
  }

  __shared__ float xch[2][2][WAVES][EX];
  __shared__ int sxch[2][WAVES][EX];
  // neighbour codes: bit0 L, bit1 R, bit2 B, bit3 T replaced by the centre
  // (solid or outside the grid: common.sh clamp == Poisson.comp obstacle
  // rule); bit4 the cell itself is not fluid.
#pragma unroll
  for (int c = 0; c < COLS; ++c) {
    sxch[0][wave][COLS * lane + c] = sol[0][c];
    sxch[1][wave][COLS * lane + c] = sol[ROWS - 1][c];
  }
  __syncthreads();
  int cd[ROWS][COLS];
  bool any_nbr = false;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int sl = ilane_from_left(sol[r][COLS - 1]);
    const int sr = ilane_from_right(sol[r][0]);
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const int L = c > 0 ? sol[r][c - 1] : sl;
      const int R = c < COLS - 1 ? sol[r][c + 1] : sr;
      const int B = r > 0 ? sol[r - 1][c]
                          : (wave > 0 ? sxch[1][wave - 1][COLS * lane + c] : 1);
      const int T = r < ROWS - 1 ? sol[r + 1][c]
                                 : (wave < WAVES - 1 ? sxch[0][wave + 1][COLS * lane + c] : 1);
      const int code = L | (R << 1) | (B << 2) | (T << 3) | (sol[r][c] << 4);
      cd[r][c] = code;
      any_nbr |= (code & 0xF) != 0;
    }
  }
  const bool special = __any(any_nbr);

  double sr = 0.0, srr = 0.0;
  const int trow0 = wave * ROWS;  // tile row of p[0][*]
  const int tcol0 = COLS * lane;  // tile column of p[*][0]

  for (int s = 0; s < k; ++s) {
    const int buf = s & 1;
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      xch[buf][0][wave][tcol0 + c] = p[0][c];
      xch[buf][1][wave][tcol0 + c] = p[ROWS - 1][c];
    }
    __syncthreads();
    float prev[COLS], above[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      prev[c] = wave > 0 ? xch[buf][1][wave - 1][tcol0 + c] : p[0][c];
      above[c] = wave < WAVES - 1 ? xch[buf][0][wave + 1][tcol0 + c] : p[ROWS - 1][c];
    }
    const bool last = (s == k - 1) && red != nullptr;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const float lft = lane_from_left(p[r][COLS - 1]);
      const float rgt = lane_from_right(p[r][0]);
      float np[COLS];
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const float pc = p[r][c];
        float x1 = c > 0 ? p[r][c - 1] : lft;
        float x2 = c < COLS - 1 ? p[r][c + 1] : rgt;
        float y1 = prev[c];
        float y2 = r < ROWS - 1 ? p[r + 1][c] : above[c];
        if (special) {
          const int code = cd[r][c];
          x1 = (code & 1) ? pc : x1;
          x2 = (code & 2) ? pc : x2;
          y1 = (code & 4) ? pc : y1;
          y2 = (code & 8) ? pc : y2;
        }
        np[c] = ((((x1 + x2) + y1) + y2) - bb[r][c]) * 0.25f;
        if (last) {
          const int trow = trow0 + r, tcol = tcol0 + c;
          if (tcol >= k && tcol < EX - k && trow >= k && trow < EY - k &&
              (cd[r][c] & 16) == 0) {
            const double rr = (double)bb[r][c] -
                              (((((double)x1 + (double)x2) + (double)y1) +
                                (double)y2) -
                               4.0 * (double)pc);
            sr += rr;
            srr += rr * rr;
          }
        }
      }
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        prev[c] = p[r][c];
        p[r][c] = np[c];
      }
    }
  }

#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int trow = trow0 + r;
    const int gy = gy0 + r;
    if (trow < k || trow >= EY - k || gy < 0 || gy >= H) continue;
    if (vec) {
      if (tcol0 >= k && tcol0 < EX - k && gx0 >= 0 && gx0 < W)
        *reinterpret_cast<float2*>(pout + gy * W + gx0) =
            make_float2(p[r][0], p[r][COLS - 1]);
    } else {
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gx = gx0 + c, tcol = tcol0 + c;
        if (tcol >= k && tcol < EX - k && gx >= 0 && gx < W)
          pout[gy * W + gx] = p[r][c];
      }
    }
  }

  if (red != nullptr) {
    __shared__ double part[WAVES][2];
    sr = wave_sum(sr);
    srr = wave_sum(srr);
    if (lane == 0) {
      part[wave][0] = sr;
      part[wave][1] = srr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0.0, c2 = 0.0;
      for (int w = 0; w < WAVES; ++w) {
        a += part[w][0];
        c2 += part[w][1];
      }
      const int bid = blockIdx.y * gridDim.x + blockIdx.x;
      red[2 * bid + 0] = a;
      red[2 * bid + 1] = c2;
    }
  }
}

// Exact FP64 residual of p (or of a zero guess when p == nullptr) plus the
// right-hand-side sums, per 64x4 block: red[5*bid + {sum r, sum r^2, sum b,
// sum b^2, fluid count}] over fluid cells.
__global__ void __launch_bounds__(256)
    k_sums(const float* __restrict__ p, const float* __restrict__ b,
           const uint8_t* __restrict__ mask, int W, int H,
           double* __restrict__ red) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  double v[5] = {0, 0, 0, 0, 0};
  if (x < W && y < H) {
    const int pos = y * W + x;
    if (mask[pos] == 0) {
      const double bd = (double)b[pos];
      double rr = bd;
      if (p != nullptr) {
        const int iL = y * W + max(x - 1, 0), iR = y * W + min(x + 1, W - 1);
        const int iB = max(y - 1, 0) * W + x, iT = min(y + 1, H - 1) * W + x;
        const float pc = p[pos];
        const float x1 = mask[iL] ? pc : p[iL];
        const float x2 = mask[iR] ? pc : p[iR];
        const float y1 = mask[iB] ? pc : p[iB];
        const float y2 = mask[iT] ? pc : p[iT];
        rr = bd - (((((double)x1 + (double)x2) + (double)y1) + (double)y2) -
                   4.0 * (double)pc);
      }
      v[0] = rr;
      v[1] = rr * rr;
      v[2] = bd;
      v[3] = bd * bd;
      v[4] = 1.0;
    }
  }
  __shared__ double part[4][5];
  const int lane = threadIdx.x, wave = threadIdx.y;
#pragma unroll
  for (int j = 0; j < 5; ++j) v[j] = wave_sum(v[j]);
  if (lane == 0)
    for (int j = 0; j < 5; ++j) part[wave][j] = v[j];
  __syncthreads();
  if (lane == 0 && wave == 0) {
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    for (int j = 0; j < 5; ++j)
      red[5 * bid + j] = ((part[0][j] + part[1][j]) + part[2][j]) + part[3][j];
  }
}

__device__ double rel_from(const double* s) {
  const double n = s[4];
  if (n <= 0.0) return 0.0;
  const double num = fmax(s[1] - s[0] * s[0] / n, 0.0);
  const double den = fmax(s[3] - s[2] * s[2] / n, 0.0);
  if (den == 0.0) return num == 0.0 ? 0.0 : __builtin_inf();
  return sqrt(num / den);
}

// Deterministic single-block reduction of `nblocks` partial records of
// `width` doubles into sums[off .. off+width).  Then sums[5] = relative
// residual.  If state != nullptr (tolerance-driven solve): on the first
// time rel <= tol it sets state[0] = 1 (done), state[1] = sweeps_now and
// state[2] = final_idx; state[3] always records sweeps_now.
__global__ void __launch_bounds__(256)
    k_reduce(const double* __restrict__ part, int nblocks, int width, int off,
             double* __restrict__ sums, double tol, int* __restrict__ state,
             int sweeps_now, int final_idx) {
  __shared__ double acc[256];
  for (int j = 0; j < width; ++j) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += 256) a += part[i * width + j];
    acc[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) acc[threadIdx.x] += acc[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) sums[off + j] = acc[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rel = rel_from(sums);
    sums[5] = rel;
    if (state != nullptr) {
      if (state[0] == 0 && rel <= tol) {
        state[0] = 1;
        state[1] = sweeps_now;
        state[2] = final_idx;
      }
      state[3] = sweeps_now;
    }
  }
}

// Tolerance-driven solve only: remove the fluid-cell mean m = sums[2]/sums[4]
// from b and restart p from the first sweep of the centred system,
// p = (0 - b')*0.25.  The operator's row sums are zero, so a constant in b
// only drifts p by -m/4 per sweep on the fluid cells and never changes the
// fluid velocity or the mean-free residual; centring keeps |p| (and with it
// the float32 round-off floor of the residual) from growing with the sweep
// count.
__global__ void __launch_bounds__(256)
    k_center_rhs(float* __restrict__ b, float* __restrict__ p,
                 const uint8_t* __restrict__ mask, int n,
                 const double* __restrict__ sums) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || mask[i] != 0) return;
  const double cnt = sums[4];
  const float m = cnt > 0.0 ? (float)(sums[2] / cnt) : 0.0f;
  const float bc = b[i] - m;
  b[i] = bc;
  if (p != nullptr) p[i] = (0.0f - bc) * 0.25f;
}

template <int ROWS, int WAVES, int COLS>
hipError_t launch_tb(const float* pin, const float* b, const uint8_t* mask,
                     float* pout, int W, int H, int k, double* red,
                     const int* done, hipStream_t s) {
  constexpr int EX = 64 * COLS, EY = ROWS * WAVES;
  const int VX = EX - 2 * k, VY = EY - 2 * k;
  dim3 grid((W + VX - 1) / VX, (H + VY - 1) / VY, 1);
  k_jacobi_tb<ROWS, WAVES, COLS><<<grid, WAVES * 64, 0, s>>>(
      pin, b, mask, pout, W, H, k, red, done);
  return hipGetLastError();
}

}  // namespace

// Instantiated tile shapes (rows per wave, waves per block, columns per
// lane).  The table is the autotuner's search space.
#define NX_TB_SHAPES(X) \
  X(4, 8, 1)            \
  X(8, 8, 1)            \
  X(8, 16, 1)           \
  X(16, 8, 1)           \
  X(2, 8, 2)            \
  X(2, 16, 2)           \
  X(4, 4, 2)            \
  X(4, 8, 2)            \
  X(4, 16, 2)           \
  X(8, 4, 2)            \
  X(8, 8, 2)

bool jacobi_cfg_supported(JacobiCfg c) {
  bool shape = false;
#define NX_TB_CHECK(R, WV, CL) \
  shape |= (c.rows == R && c.waves == WV && c.cols == CL);
  NX_TB_SHAPES(NX_TB_CHECK)
#undef NX_TB_CHECK
  if (!shape) return false;
  const int EX = 64 * c.cols, EY = c.rows * c.waves;
  return c.k >= 1 && EX - 2 * c.k >= 8 && EY - 2 * c.k >= 8;
}

int jacobi_launch_blocks(int W, int H, JacobiCfg c) {
  const int VX = 64 * c.cols - 2 * c.k, VY = c.rows * c.waves - 2 * c.k;
  return ((W + VX - 1) / VX) * ((H + VY - 1) / VY);
}

hipError_t launch_jacobi_tb(const float* pin, const float* b,
                            const uint8_t* mask, float* pout, int W, int H,
                            JacobiCfg c, double* red, const int* done,
                            hipStream_t s) {
  if (!jacobi_cfg_supported(c)) return hipErrorInvalidValue;
#define NX_TB_LAUNCH(R, WV, CL)                                              \
  if (c.rows == R && c.waves == WV && c.cols == CL)                          \
    return launch_tb<R, WV, CL>(pin, b, mask, pout, W, H, c.k, red, done, s);
  NX_TB_SHAPES(NX_TB_LAUNCH)
#undef NX_TB_LAUNCH
  return hipErrorInvalidValue;
}

hipError_t launch_center_rhs(float* b, float* p, const uint8_t* mask, int n,
                             const double* sums, hipStream_t s) {
  k_center_rhs<<<(n + 255) / 256, 256, 0, s>>>(b, p, mask, n, sums);
  return hipGetLastError();
}

int residual_blocks(int W, int H) { return ((W + 63) / 64) * ((H + 3) / 4); }

hipError_t launch_sums(const float* p, const float* b, const uint8_t* mask,
                       int W, int H, double* red, hipStream_t s) {
  dim3 grid((W + 63) / 64, (H + 3) / 4, 1);
  k_sums<<<grid, dim3(64, 4), 0, s>>>(p, b, mask, W, H, red);
  return hipGetLastError();
}

hipError_t launch_reduce(const double* part, int nblocks, int width, int off,
                         double* sums, double tol, int* state, int sweeps_now,
                         int final_idx, hipStream_t s) {
  k_reduce<<<1, 256, 0, s>>>(part, nblocks, width, off, sums, tol, state,
                             sweeps_now, final_idx);
  return hipGetLastError();
}

}  // namespace nx
