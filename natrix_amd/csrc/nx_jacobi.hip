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
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"

namespace nx {

namespace {

using jt::wave_sum;

template <int ROWS, int WAVES, int COLS>
__global__ void __launch_bounds__(WAVES * 64)
    __attribute__((amdgpu_waves_per_eu(jt::jacobi_waves_per_eu<ROWS * COLS>())))
    k_jacobi_tb(const float* __restrict__ pin, const float* __restrict__ b,
                const uint8_t* __restrict__ mask, float* __restrict__ pout,
                int W, int H, int k, double* __restrict__ red,
                const int* __restrict__ done) {
  if (done != nullptr && *done != 0) return;  // block-uniform early exit
  // 1D grid of gx * gy tiles in XCD order; the partial-sum record is the
  // tile's row-major index
  const int gx = (W + 64 * COLS - 2 * k - 1) / (64 * COLS - 2 * k);
  const int gy = (H + ROWS * WAVES - 2 * k - 1) / (ROWS * WAVES - 2 * k);
  const int t = xcd_tile<XCD_RUNS>((int)blockIdx.x, gx * gy);
  jt::jacobi_tb_tile<ROWS, WAVES, COLS>(t % gx, t / gx, pin, b, mask, pout, W, H, k, red,
                                        t);
}

// Exact FP64 residual of p (or of a zero guess when p == nullptr) plus the
// right-hand-side sums, per 64x4 block: red[5*bid + {sum r, sum r^2, sum b,
// sum b^2, fluid count}] over fluid cells.
__global__ void __launch_bounds__(256)
    k_sums(const float* __restrict__ p, const float* __restrict__ b,
           const uint8_t* __restrict__ mask, int W, int H,
           double* __restrict__ red) {
  // a block covers 64 x RES_ROWS cells: wave w takes rows w, w+4, ... so
  // every load instruction reads one contiguous row segment
  const int x = blockIdx.x * 64 + threadIdx.x;
  double v[5] = {0, 0, 0, 0, 0};
#pragma unroll 2
  for (int r = 0; r < RES_ROWS / 4; ++r) {
    const int y = blockIdx.y * RES_ROWS + threadIdx.y + 4 * r;
    if (x >= W || y >= H) continue;
    const int pos = y * W + x;
    if (mask[pos] != 0) continue;
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
    v[0] += rr;
    v[1] += rr * rr;
    v[2] += bd;
    v[3] += bd * bd;
    v[4] += 1.0;
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
__global__ void __launch_bounds__(1024)
    k_reduce(const double* __restrict__ part, int nblocks, int width, int off,
             double* __restrict__ sums, double tol, int* __restrict__ state,
             int sweeps_now, int final_idx, double* __restrict__ mirror,
             const double* __restrict__ extra) {
  // one pass over the records: thread t adds records t, t + 1024, ... in
  // order, then the 16 wave sums are added in wave order (deterministic)
  double a[5] = {0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < nblocks; i += 1024)
#pragma unroll
    for (int j = 0; j < 5; ++j)
      if (j < width) a[j] += part[i * width + j];
  __shared__ double acc[16][5];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 5; ++j) a[j] = wave_sum(a[j]);
  if (lane == 0)
    for (int j = 0; j < 5; ++j) acc[wave][j] = a[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < width; ++j) {
      double t = 0.0;
      for (int w = 0; w < 16; ++w) t += acc[w][j];
      sums[off + j] = t;
    }
    const double rel = rel_from(sums);
    sums[5] = rel;
    if (mirror != nullptr) {
      mirror[0] = rel;
      mirror[1] = extra != nullptr ? extra[0] : 0.0;
    }
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
  const dim3 grid(((W + VX - 1) / VX) * ((H + VY - 1) / VY), 1, 1);
  k_jacobi_tb<ROWS, WAVES, COLS><<<grid, WAVES * 64, 0, s>>>(
      pin, b, mask, pout, W, H, k, red, done);
  return hipGetLastError();
}

}  // namespace

// Instantiated tile shapes (rows per wave, waves per block, columns per
// lane).  The table is the autotuner's search space.
#define NX_TB_SHAPES(X) \
  X(4, 8, 1)            \
  X(4, 16, 1)           \
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
  if (c.seg != 0) return c.seg > 0 && stream_cfg_supported(c);
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
  if (c.seg > 0) return stream_launch_blocks(W, H, c);
  const int VX = 64 * c.cols - 2 * c.k, VY = c.rows * c.waves - 2 * c.k;
  return ((W + VX - 1) / VX) * ((H + VY - 1) / VY);
}

hipError_t launch_jacobi_tb(const float* pin, const float* b,
                            const uint8_t* mask, float* pout, int W, int H,
                            JacobiCfg c, double* red, const int* done,
                            hipStream_t s) {
  if (!jacobi_cfg_supported(c)) return hipErrorInvalidValue;
  if (c.seg > 0) return launch_jacobi_stream(pin, b, mask, pout, W, H, c, red, done, s);
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

int residual_blocks(int W, int H) {
  return ((W + 63) / 64) * ((H + RES_ROWS - 1) / RES_ROWS);
}

hipError_t launch_sums(const float* p, const float* b, const uint8_t* mask,
                       int W, int H, double* red, hipStream_t s) {
  dim3 grid((W + 63) / 64, (H + RES_ROWS - 1) / RES_ROWS, 1);
  k_sums<<<grid, dim3(64, 4), 0, s>>>(p, b, mask, W, H, red);
  return hipGetLastError();
}

hipError_t launch_reduce(const double* part, int nblocks, int width, int off,
                         double* sums, double tol, int* state, int sweeps_now,
                         int final_idx, hipStream_t s, double* mirror,
                         const double* extra) {
  k_reduce<<<1, 1024, 0, s>>>(part, nblocks, width, off, sums, tol, state,
                              sweeps_now, final_idx, mirror, extra);
  return hipGetLastError();
}

}  // namespace nx
