// Temporal-blocked Jacobi tile (shared by the single-grid and batched
// kernels); see nx_jacobi.hip for the scheme.
#pragma once
#include <type_traits>

#include "nx_common.h"

namespace nx {
namespace jt {

// Waves per SIMD asked of the compiler for a tile kernel of CELLS cells
// per lane (amdgpu_waves_per_eu; 1 leaves it free).  Left free, a Jacobi
// tile of 8 cells per lane takes 62 VGPRs and 106 SGPRs, which caps it at 7
// waves per SIMD, 3 blocks of 8 waves per CU; asked for 8, the compiler
// keeps the SGPRs that do not fit in VGPR lanes and a 4th block fits.  A
// Chebyshev tile of 8 cells per lane takes 83 VGPRs (5 waves, 2 blocks of 8
// waves); asked for 6 it fits in 80 with one spilled VGPR, and a 3rd block
// fits.  Its residual variant and the larger tiles would spill heavily, so
// they stay free.  Tiles of 4 cells per lane ask for 8, so that two blocks
// of 16 waves fit.
template <int CELLS>
constexpr int jacobi_waves_per_eu() { return CELLS <= 8 ? 8 : 1; }
template <int CELLS, bool RES>
constexpr int cheb_waves_per_eu() { return CELLS <= 4 ? 8 : (CELLS == 8 && !RES ? 6 : 1); }

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

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

__device__ __forceinline__ int ilane_from_left(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}

__device__ __forceinline__ int ilane_from_right(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false);
}

// The same shifts with the edge lane reading 0 (bound_ctrl): no old value
// to keep, so the compiler folds the shift into the instruction that
// consumes it as a DPP source operand (v_add_f32_dpp), at no cost of its
// own.  For callers whose edge lanes hold cells outside the valid region.
__device__ __forceinline__ float lane_shr0(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float lane_shl0(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// ROWS x COLS cells per lane: lane `l` of wave `w` owns tile columns
// [COLS*l, COLS*l + COLS) and tile rows [ROWS*w, ROWS*w + ROWS).
// k sweeps on tile (bx, by) of the (VX x VY)-interior tiling; `rid` is the
// partial-sum record this tile writes when red != nullptr.  Uses
// __syncthreads(): every thread of the block must call it.
// p and b are stored in format S and held in float32 registers across the k
// sweeps; p is rounded to S once, at the store.
// CORR (residual correction): the first of the k passes computes the float32
// residual r = b - A p of the stored iterate instead of a sweep, the other
// k-1 passes are Jacobi sweeps of A e = r from e = 0, and the tile writes
// the float32 pressure p + e to pout32 (pout is unused).
template <int ROWS, int WAVES, int COLS, class S = F32S, bool CORR = false>
__device__ __forceinline__ void jacobi_tb_tile(
    int bx, int by, const typename S::s1* __restrict__ pin,
    const typename S::s1* __restrict__ b, const uint8_t* __restrict__ mask,
    typename S::s1* __restrict__ pout, int W, int H, int k,
    double* __restrict__ red, int rid, float* __restrict__ pout32 = nullptr) {
  constexpr int EX = 64 * COLS, EY = ROWS * WAVES;
  const int lane = threadIdx.x & 63;
  // wave-uniform: a scalar register, so branches on it are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int VX = EX - 2 * k, VY = EY - 2 * k;
  const int gx0 = bx * VX - k + COLS * lane;
  const int gy0 = by * VY - k + wave * ROWS;
  // float2 loads/stores when every lane's column pair is aligned and either
  // wholly inside or wholly outside the grid and the valid band
  const bool vec = COLS == 2 && (W % 2 == 0) && (k % 2 == 0);

  float p[ROWS][COLS], bb[ROWS][COLS];
  int sol[ROWS][COLS];  // cell is solid or outside the grid
  // Every load is unconditional, from a clamped in-grid address, and all of
  // a lane's loads are issued before the first is used; cells outside the
  // grid are replaced afterwards.
  if (vec) {
    const int cx = min(max(gx0, 0), W - 2);  // even, like gx0
    const bool col_in = gx0 >= 0 && gx0 < W;
    float2 lp[ROWS], lb[ROWS];
    uint16_t lm[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int idx = min(max(gy0 + r, 0), H - 1) * W + cx;
      lp[r] = S::ldp(pin, idx);
      lb[r] = S::ldp(b, idx);
      lm[r] = *reinterpret_cast<const uint16_t*>(mask + idx);
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const bool in = col_in && gy0 + r >= 0 && gy0 + r < H;
      p[r][0] = in ? lp[r].x : 0.0f;
      p[r][COLS - 1] = in ? lp[r].y : 0.0f;
      bb[r][0] = in ? lb[r].x : 0.0f;
      bb[r][COLS - 1] = in ? lb[r].y : 0.0f;
      sol[r][0] = in ? (lm[r] & 0xFF) != 0 : 1;
      sol[r][COLS - 1] = in ? (lm[r] >> 8) != 0 : 1;
    }
  } else {
    float lp[ROWS][COLS], lb[ROWS][COLS];
    uint8_t lm[ROWS][COLS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int idx = min(max(gy0 + r, 0), H - 1) * W + min(max(gx0 + c, 0), W - 1);
        lp[r][c] = S::ld(pin, idx);
        lb[r][c] = S::ld(b, idx);
        lm[r][c] = mask[idx];
      }
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gy = gy0 + r, gx = gx0 + c;
        const bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
        p[r][c] = in ? lp[r][c] : 0.0f;
        bb[r][c] = in ? lb[r][c] : 0.0f;
        sol[r][c] = in ? (lm[r][c] != 0) : 1;
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
      // the rows beyond the tile's first and last count as fluid: their
      // cells are outside the valid cone, and the waves holding them then
      // take the fast path too
      const int B = r > 0 ? sol[r - 1][c]
                          : (wave > 0 ? sxch[1][wave - 1][COLS * lane + c] : 0);
      const int T = r < ROWS - 1 ? sol[r + 1][c]
                                 : (wave < WAVES - 1 ? sxch[0][wave + 1][COLS * lane + c] : 0);
      const int code = L | (R << 1) | (B << 2) | (T << 3) | (sol[r][c] << 4);
      cd[r][c] = code;
      any_nbr |= (code & 0xF) != 0;
    }
  }
  const bool special = __any(any_nbr);

  double sr = 0.0, srr = 0.0;
  const int trow0 = wave * ROWS;  // tile row of p[0][*]
  const int tcol0 = COLS * lane;  // tile column of p[*][0]

  float pst[CORR ? ROWS : 1][CORR ? COLS : 1];  // stored iterate (CORR)
  int s0 = 0;
  if (CORR) {
    // pass 0: r = b - ((((x1 + x2) + y1) + y2) - 4 p), float32
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      xch[0][0][wave][tcol0 + c] = p[0][c];
      xch[0][1][wave][tcol0 + c] = p[ROWS - 1][c];
    }
    __syncthreads();
    float prev[COLS], above[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      prev[c] = wave > 0 ? xch[0][1][wave - 1][tcol0 + c] : p[0][c];
      above[c] = wave < WAVES - 1 ? xch[0][0][wave + 1][tcol0 + c] : p[ROWS - 1][c];
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const float lft = lane_from_left(p[r][COLS - 1]);
      const float rgt = lane_from_right(p[r][0]);
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const float pc = p[r][c];
        const int code = cd[r][c];
        const float x1 = (code & 1) ? pc : (c > 0 ? p[r][c - 1] : lft);
        const float x2 = (code & 2) ? pc : (c < COLS - 1 ? p[r][c + 1] : rgt);
        const float y1 = (code & 4) ? pc : prev[c];
        const float y2 = (code & 8) ? pc : (r < ROWS - 1 ? p[r + 1][c] : above[c]);
        bb[r][c] = bb[r][c] - ((((x1 + x2) + y1) + y2) - 4.0f * pc);
      }
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        prev[c] = p[r][c];
        pst[CORR ? r : 0][CORR ? c : 0] = p[r][c];
        p[r][c] = 0.0f;
      }
    }
    s0 = 1;
  }

  // the rows above and below a wave's strip come from its neighbours in the
  // block; the first and last waves read their own published rows instead
  // (tile rows outside every sweep's valid cone, so any finite value
  // serves), which keeps the LDS reads unconditional
  const int wave_up = wave > 0 ? wave - 1 : 0;
  const int wave_dn = wave < WAVES - 1 ? wave + 1 : WAVES - 1;
  auto exchange = [&](int buf, float(&prev)[COLS], float(&above)[COLS]) {
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      xch[buf][0][wave][tcol0 + c] = p[0][c];
      xch[buf][1][wave][tcol0 + c] = p[ROWS - 1][c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      prev[c] = xch[buf][1][wave_up][tcol0 + c];
      above[c] = xch[buf][0][wave_dn][tcol0 + c];
    }
  };

  // Sweeps in two phases: the neighbour sums of every cell, then the
  // update in place (p is dead once the sums are formed, so no register
  // copies).  SP: with neighbour replacement (a solid or out-of-grid
  // neighbour somewhere in the wave); wave-uniform, one scalar branch.
  // The last sweep of a launch that accumulates the residual is peeled.
  int s = s0;
  const int s_fast = red != nullptr ? k - 1 : k;
  auto sweeps = [&](auto spc) {
    constexpr bool SP = decltype(spc)::value;
    for (; s < s_fast; ++s) {
      float prev[COLS], above[COLS];
      exchange(s & 1, prev, above);
      float sum[ROWS][COLS];
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const float pc = p[r][c];
          // lane 0's left and lane 63's right neighbour read 0: tile
          // columns outside the valid cone
          float x1 = c > 0 ? p[r][c - 1] : lane_shr0(p[r][COLS - 1]);
          float x2 = c < COLS - 1 ? p[r][c + 1] : lane_shl0(p[r][0]);
          float y1 = r > 0 ? p[r - 1][c] : prev[c];
          float y2 = r < ROWS - 1 ? p[r + 1][c] : above[c];
          if (SP) {
            const int code = cd[r][c];
            x1 = (code & 1) ? pc : x1;
            x2 = (code & 2) ? pc : x2;
            y1 = (code & 4) ? pc : y1;
            y2 = (code & 8) ? pc : y2;
          }
          sum[r][c] = ((x1 + x2) + y1) + y2;
        }
      }
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
#pragma unroll
        for (int c = 0; c < COLS; ++c) p[r][c] = (sum[r][c] - bb[r][c]) * 0.25f;
    }
  };
  if (special)
    sweeps(std::true_type());
  else
    sweeps(std::false_type());

  if (red != nullptr && s < k) {
    // the last sweep, accumulating the FP64 residual of its input iterate
    float prev[COLS], above[COLS];
    exchange(s & 1, prev, above);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const float lft = lane_from_left(p[r][COLS - 1]);
      const float rgt = lane_from_right(p[r][0]);
      float np[COLS];
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const float pc = p[r][c];
        const int code = cd[r][c];
        const float x1 = (code & 1) ? pc : (c > 0 ? p[r][c - 1] : lft);
        const float x2 = (code & 2) ? pc : (c < COLS - 1 ? p[r][c + 1] : rgt);
        const float y1 = (code & 4) ? pc : prev[c];
        const float y2 = (code & 8) ? pc : (r < ROWS - 1 ? p[r + 1][c] : above[c]);
        np[c] = ((((x1 + x2) + y1) + y2) - bb[r][c]) * 0.25f;
        const int trow = trow0 + r, tcol = tcol0 + c;
        if (tcol >= k && tcol < EX - k && trow >= k && trow < EY - k && (code & 16) == 0) {
          const double rr =
              (double)bb[r][c] -
              (((((double)x1 + (double)x2) + (double)y1) + (double)y2) - 4.0 * (double)pc);
          sr += rr;
          srr += rr * rr;
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
    if (CORR) {
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gx = gx0 + c, tcol = tcol0 + c;
        if (tcol >= k && tcol < EX - k && gx >= 0 && gx < W)
          pout32[gy * W + gx] = pst[CORR ? r : 0][CORR ? c : 0] + p[r][c];
      }
    } else if (vec) {
      if (tcol0 >= k && tcol0 < EX - k && gx0 >= 0 && gx0 < W)
        S::stp(pout, gy * W + gx0, make_float2(p[r][0], p[r][COLS - 1]));
    } else {
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gx = gx0 + c, tcol = tcol0 + c;
        if (tcol >= k && tcol < EX - k && gx >= 0 && gx < W)
          S::st(pout, gy * W + gx, p[r][c]);
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
      red[2 * rid + 0] = a;
      red[2 * rid + 1] = c2;
    }
  }
}

}  // namespace jt
}  // namespace nx
