// Streaming temporal-blocked stencil kernels: the fine Chebyshev smoother
// and the Jacobi sweeps, with no vertical halo re-reads.
//
// The tile kernels (nx_cheb.hip, nx_jacobi_tile.h) load a (64 COLS) x
// (ROWS WAVES) tile, run k passes and write back the interior their
// dependency cone kept, so both the horizontal and the vertical k-cell halo
// are re-read by the neighbouring tiles.  Here a workgroup owns a strip of
// EX = 64 COLS columns over a vertical segment [Y0, Y1) of rows, and walks
// the segment in chunks of S = ROWS WAVES rows.  Within a chunk the passes
// form a parallelogram: after pass s a register row holds its level-0 row
// moved up by s + 1 grid rows, so pass s + 1 at register row q reads the
// previous level at register rows q - 2, q - 1 (centre) and q.  The two rows
// above a chunk at every level are the last rows the previous chunk held at
// that level: wave WAVES - 1 leaves them in a per-pass LDS carry for wave 0
// of the next chunk.  Vertically only 2k rows per segment are loaded twice
// (the warm-up of its first chunk), horizontally the k-column halo of the
// strip remains; each segment writes Y1 - Y0 rows of EX - 2k columns.
//
// Every cell evaluates the same expressions in the same order as one pass of
// the tile kernels (a solid or out-of-grid neighbour reads the centre value:
// shader.Poisson.comp:26-37 with the clamp of shader.common.sh:14-17), so
// the result is bitwise that of the tile kernels and of single passes.
//
// Geometry of block (bx, by): lane l of wave w owns columns
// gx0 = bx (EX - 2k) - k + COLS l ... + COLS - 1; Y0 = by seg, Y1 =
// min(Y0 + seg, H); chunk c loads level-0 rows A + c S + w ROWS + i with
// A = Y0 - k; after pass s register row (w, i) holds grid row
// A + c S + w ROWS + i - (s + 1).  Rows above Y0 - k are never loaded: the
// first chunk's carry is zero with solid flags, which only reaches rows
// above Y0 at level k (row Y0 - k + s at pass s is the first valid one).
// Chunks per segment: ceil((Y1 - Y0 + 2k) / S).
#include <type_traits>

#include "../../include/natrix.h"
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"

namespace nx {

namespace {

using jt::ilane_from_left;
using jt::ilane_from_right;
using jt::lane_from_left;
using jt::lane_from_right;

// LDS exchange of one pass: the stencil operand's last two rows, two other
// per-cell values of the last row, and the solid flags of the last two rows
// (bit 0: row ROWS - 2, bit 1: row ROWS - 1).
template <int EX, int NV = 4>
struct Pub {
  float v[NV][EX];
  uint8_t s[EX];
};

enum PassKind { P_SETUP = 0, P_ITER = 1, P_RES = 2 };

// One Chebyshev pass over a wave's register rows, descending, in place.
// V is the stencil operand (d for iterations, x for setup and residual),
// (A, B) the other two arrays ((x, r) for iterations, (r, d) otherwise).
// p0..p3 / ps: the previous wave's (or the carry's) last rows.
template <int KIND, int ROWS, int COLS>
__device__ __forceinline__ void cheb_pass(float (&X)[ROWS][COLS], float (&R)[ROWS][COLS],
                                          float (&D)[ROWS][COLS], int (&SL)[ROWS][COLS],
                                          const float (&p0)[COLS], const float (&p1)[COLS],
                                          const float (&p2)[COLS], const float (&p3)[COLS],
                                          const int (&ps)[COLS], float theta, float cdd,
                                          float crr, int row_new0, int gx0, int tcol0, int k,
                                          int Y0, int Y1, int W, int EX,
                                          const float (&BR)[ROWS][COLS], double (&acc)[5],
                                          bool special) {
  float(&V)[ROWS][COLS] = KIND == P_ITER ? D : X;
  float(&Aa)[ROWS][COLS] = KIND == P_ITER ? X : R;
  float(&Bb)[ROWS][COLS] = KIND == P_ITER ? R : D;
#pragma unroll
  for (int i = ROWS - 1; i >= 0; --i) {
    float vc[COLS], va[COLS], vb[COLS], ac[COLS], bc[COLS];
    int sc[COLS], sa[COLS], sb[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      vc[c] = i >= 1 ? V[i - 1][c] : p1[c];
      va[c] = i >= 2 ? V[i - 2][c] : (i == 1 ? p1[c] : p0[c]);
      vb[c] = V[i][c];
      ac[c] = i >= 1 ? Aa[i - 1][c] : p2[c];
      bc[c] = i >= 1 ? Bb[i - 1][c] : p3[c];
      sc[c] = i >= 1 ? SL[i - 1][c] : (ps[c] >> 1) & 1;
      sa[c] = i >= 2 ? SL[i - 2][c] : (i == 1 ? (ps[c] >> 1) & 1 : ps[c] & 1);
      sb[c] = SL[i][c];
    }
    const float lft = lane_from_left(vc[COLS - 1]);
    const float rgt = lane_from_right(vc[0]);
    int slf = 0, srg = 0;
    if (special) {
      slf = ilane_from_left(sc[COLS - 1]);
      srg = ilane_from_right(sc[0]);
    }
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const float v = vc[c];
      float x1 = c > 0 ? vc[c - 1] : lft;
      float x2 = c < COLS - 1 ? vc[c + 1] : rgt;
      float y1 = va[c];
      float y2 = vb[c];
      if (special) {
        const int s1 = c > 0 ? sc[c - 1] : slf;
        const int s2 = c < COLS - 1 ? sc[c + 1] : srg;
        x1 = s1 ? v : x1;
        x2 = s2 ? v : x2;
        y1 = sa[c] ? v : y1;
        y2 = sb[c] ? v : y2;
      }
      if (KIND == P_SETUP) {
        const float lv = (((x1 + x2) + y1) + y2) - 4.0f * v;
        const float rn = lv - ac[c];  // r holds b before the setup
        X[i][c] = v;
        R[i][c] = rn;
        D[i][c] = rn / theta;
      } else if (KIND == P_ITER) {
        const float lv = (((x1 + x2) + y1) + y2) - 4.0f * v;
        const float rn = bc[c] + lv;
        X[i][c] = ac[c] + v;
        R[i][c] = rn;
        D[i][c] = cdd * v + crr * rn;
      } else {
        const int gy = row_new0 + i, gx = gx0 + c, tcol = tcol0 + c;
        if (gy >= Y0 && gy < Y1 && tcol >= k && tcol < EX - k && gx >= 0 && gx < W &&
            !sc[c]) {
          const double bd = (double)BR[i][c];  // b of the output row, loaded with the chunk
          const double rr =
              bd - (((((double)x1 + (double)x2) + (double)y1) + (double)y2) - 4.0 * (double)v);
          acc[0] += rr;
          acc[1] += rr * rr;
          acc[2] += bd;
          acc[3] += bd * bd;
          acc[4] += 1.0;
        }
        X[i][c] = v;
        R[i][c] = ac[c];
        D[i][c] = bc[c];
      }
      SL[i][c] = sc[c];
    }
  }
}

template <int ROWS, int WAVES, int COLS, bool RES>
__global__ void __launch_bounds__(WAVES * 64)
    k_cheb_stream(const float* __restrict__ xin, const float* __restrict__ rin,
                  const float* __restrict__ din, const float* __restrict__ b,
                  const uint8_t* __restrict__ mask, float* __restrict__ xout,
                  float* __restrict__ rout, float* __restrict__ dout, int W, int H, int k,
                  int seg, int first, float theta, ChebCoef cf, double* __restrict__ red,
                  const float* __restrict__ dapp, const double* __restrict__ line) {
  constexpr int EX = 64 * COLS, S = ROWS * WAVES;
  static_assert(ROWS >= 2, "the pass reads two rows above a wave's first row");
  static_assert(WAVES >= 2, "wave WAVES - 1 keeps the carry in the registers wave 0 reads it into");
  const int lane = threadIdx.x & 63;
  // the wave index is uniform across the wave: an SGPR, so branches on it are
  // scalar branches with no exec-mask merging
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tcol0 = COLS * lane;
  const int gx0 = (int)blockIdx.x * (EX - 2 * k) - k + tcol0;
  const int Y0 = (int)blockIdx.y * seg;
  const int Y1 = min(Y0 + seg, H);
  const int A = Y0 - k;
  const int nch = (Y1 - Y0 + 2 * k + S - 1) / S;
  const float a_line = dapp != nullptr ? (float)line[0] : 0.0f;
  const bool vec = COLS == 2 && (W % 2 == 0) && (k % 2 == 0);

  __shared__ Pub<EX> pub[2][WAVES];
  __shared__ Pub<EX> car[STREAM_KMAX];

  float X[ROWS][COLS], R[ROWS][COLS], D[ROWS][COLS];
  int SL[ROWS][COLS];
  // the next chunk's inputs, loaded while the current chunk's passes run
  // (its loads stay in flight across the passes' LDS barriers); with RES
  // also b of the rows the last pass writes (register rows k up)
  float Xn[ROWS][COLS], Rn[ROWS][COLS], Dn[ROWS][COLS];
  float BR[ROWS][COLS], BRn[ROWS][COLS];  // used by RES launches only
  int SLn[ROWS][COLS];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};

  auto load_chunk = [&](int ch) {
    const int row0 = A + ch * S + wave * ROWS;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i;
      const bool rin_grid = gy >= 0 && gy < H;
      // b of the row the residual pass writes from register row i
      const int gyr = gy - k;
      const bool res_row = RES && gyr >= 0 && gyr < H;
      // loads are unconditional, from a clamped in-grid address, and the
      // values of cells outside the grid are replaced afterwards (a select
      // on the loaded value keeps every load a plain global load)
      const float* rsrc = first ? b : rin;
      const float* dsrc = first ? xin : din;
      if (vec) {
        const bool in = rin_grid && gx0 >= 0 && gx0 < W;
        const int idx = in ? gy * W + gx0 : 0;
        const float2 xv = *reinterpret_cast<const float2*>(xin + idx);
        const float2 rv = *reinterpret_cast<const float2*>(rsrc + idx);
        const float2 dv = *reinterpret_cast<const float2*>(dsrc + idx);
        const uint16_t mm = *reinterpret_cast<const uint16_t*>(mask + idx);
        const bool dz = !in || first;
        Xn[i][0] = in ? xv.x : 0.0f;
        Xn[i][COLS - 1] = in ? xv.y : 0.0f;
        Rn[i][0] = in ? rv.x : 0.0f;
        Rn[i][COLS - 1] = in ? rv.y : 0.0f;
        Dn[i][0] = dz ? 0.0f : dv.x;
        Dn[i][COLS - 1] = dz ? 0.0f : dv.y;
        SLn[i][0] = in ? (mm & 0xFF) != 0 : 1;
        SLn[i][COLS - 1] = in ? (mm >> 8) != 0 : 1;
        if (RES) {
          const bool inr = res_row && gx0 >= 0 && gx0 < W;
          const float2 bv = *reinterpret_cast<const float2*>(b + (inr ? gyr * W + gx0 : 0));
          BRn[i][0] = inr ? bv.x : 0.0f;
          BRn[i][COLS - 1] = inr ? bv.y : 0.0f;
        }
      } else {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const int gx = gx0 + c;
          const bool in = rin_grid && gx >= 0 && gx < W;
          const int idx = in ? gy * W + gx : 0;
          const float xv = xin[idx], rv = rsrc[idx], dv = dsrc[idx];
          const uint8_t mv = mask[idx];
          Xn[i][c] = in ? xv : 0.0f;
          Rn[i][c] = in ? rv : 0.0f;
          Dn[i][c] = (in && !first) ? dv : 0.0f;
          SLn[i][c] = in ? (mv != 0) : 1;
          if (RES) {
            const bool inr = res_row && gx >= 0 && gx < W;
            const float bv = b[inr ? gyr * W + gx : 0];
            BRn[i][c] = inr ? bv : 0.0f;
          }
        }
      }
    }
  };

  load_chunk(0);
  for (int ch = 0; ch < nch; ++ch) {
    const int row0 = A + ch * S + wave * ROWS;  // level-0 row of register row 0
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i;
      const bool rin_grid = gy >= 0 && gy < H;
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        X[i][c] = Xn[i][c];
        R[i][c] = Rn[i][c];
        D[i][c] = Dn[i][c];
        SL[i][c] = SLn[i][c];
        BR[i][c] = RES ? BRn[i][c] : 0.0f;
      }
      // a pending line-search step x += a dapp on the fluid cells
      if (a_line != 0.0f) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const int gx = gx0 + c;
          if (rin_grid && gx >= 0 && gx < W && !SL[i][c])
            X[i][c] = X[i][c] + a_line * dapp[gy * W + gx];
        }
      }
    }
    if (ch + 1 < nch) load_chunk(ch + 1);

    // one pass s of kind KIND: exchange the rows above each wave's strip
    // (wave 0: the previous chunk's, from the carry), then update in place
    auto pass = [&](auto kindc, int s) {
      constexpr int KIND = decltype(kindc)::value;
      constexpr bool it = KIND == P_ITER;
      const int par = s & 1;
      float p0[COLS], p1[COLS], p2[COLS], p3[COLS];
      int ps[COLS];
      // wave 0 takes the previous chunk's rows at this level (read before
      // this pass's barrier; wave WAVES - 1 rewrites them after it)
      if (wave == 0) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const bool have = ch > 0;
          p0[c] = have ? car[s].v[0][tcol0 + c] : 0.0f;
          p1[c] = have ? car[s].v[1][tcol0 + c] : 0.0f;
          p2[c] = have ? car[s].v[2][tcol0 + c] : 0.0f;
          p3[c] = have ? car[s].v[3][tcol0 + c] : 0.0f;
          ps[c] = have ? car[s].s[tcol0 + c] : 3;
        }
      }
      // publish this wave's last rows of the pass's inputs
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const float v2 = it ? D[ROWS - 2][c] : X[ROWS - 2][c];
        const float v1 = it ? D[ROWS - 1][c] : X[ROWS - 1][c];
        const float a1 = it ? X[ROWS - 1][c] : R[ROWS - 1][c];
        const float b1 = it ? R[ROWS - 1][c] : D[ROWS - 1][c];
        const uint8_t sb = (uint8_t)(SL[ROWS - 2][c] | (SL[ROWS - 1][c] << 1));
        pub[par][wave].v[0][tcol0 + c] = v2;
        pub[par][wave].v[1][tcol0 + c] = v1;
        pub[par][wave].v[2][tcol0 + c] = a1;
        pub[par][wave].v[3][tcol0 + c] = b1;
        pub[par][wave].s[tcol0 + c] = sb;
        if (wave == WAVES - 1) {
          // kept for wave 0 of the next chunk; written after the barrier
          // below so wave 0's read of the old carry above comes first
          p0[c] = v2;
          p1[c] = v1;
          p2[c] = a1;
          p3[c] = b1;
          ps[c] = sb;
        }
      }
      __syncthreads();
      if (wave == WAVES - 1) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          car[s].v[0][tcol0 + c] = p0[c];
          car[s].v[1][tcol0 + c] = p1[c];
          car[s].v[2][tcol0 + c] = p2[c];
          car[s].v[3][tcol0 + c] = p3[c];
          car[s].s[tcol0 + c] = (uint8_t)ps[c];
        }
      }
      if (wave > 0) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          p0[c] = pub[par][wave - 1].v[0][tcol0 + c];
          p1[c] = pub[par][wave - 1].v[1][tcol0 + c];
          p2[c] = pub[par][wave - 1].v[2][tcol0 + c];
          p3[c] = pub[par][wave - 1].v[3][tcol0 + c];
          ps[c] = pub[par][wave - 1].s[tcol0 + c];
        }
      }
      // no solid or out-of-grid cell among the rows this pass reads in the
      // wave (its own and the two above): no neighbour is replaced.  The
      // strip's outermost lanes read garbage across the wave edge either
      // way, which stays outside the columns the launch keeps.
      int anys = ps[0] | (COLS > 1 ? ps[COLS - 1] : 0);
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c) anys |= SL[i][c];
      const bool special = __any(anys != 0);
      cheb_pass<KIND>(X, R, D, SL, p0, p1, p2, p3, ps, theta, it ? cf.cd[s] : 0.0f,
                      it ? cf.cr[s] : 0.0f, row0 - (s + 1), gx0, tcol0, k, Y0, Y1, W, EX, BR,
                      acc, special);
    };
    // the setup pass (first launch), the iterations, the residual pass (RES)
    int s = 0;
    if (first) pass(std::integral_constant<int, P_SETUP>(), s++);
    for (; s < (RES ? k - 1 : k); ++s) pass(std::integral_constant<int, P_ITER>(), s);
    if (RES) pass(std::integral_constant<int, P_RES>(), s);

    // register row i now holds grid row row0 + i - k
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i - k;
      if (gy < Y0 || gy >= Y1) continue;
      if (vec) {
        if (tcol0 < k || tcol0 >= EX - k || gx0 < 0 || gx0 >= W) continue;
        const int idx = gy * W + gx0;
        *reinterpret_cast<float2*>(xout + idx) = make_float2(X[i][0], X[i][COLS - 1]);
        if (rout) {
          *reinterpret_cast<float2*>(rout + idx) = make_float2(R[i][0], R[i][COLS - 1]);
          *reinterpret_cast<float2*>(dout + idx) = make_float2(D[i][0], D[i][COLS - 1]);
        }
        continue;
      }
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int tcol = tcol0 + c, gx = gx0 + c;
        if (tcol < k || tcol >= EX - k || gx < 0 || gx >= W) continue;
        const int idx = gy * W + gx;
        xout[idx] = X[i][c];
        if (rout) {
          rout[idx] = R[i][c];
          dout[idx] = D[i][c];
        }
      }
    }
    // the carry written in this chunk's passes is read in the next chunk's
    __syncthreads();
  }

  if (RES) {
    __shared__ double part[WAVES][5];
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] = jt::wave_sum(acc[j]);
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < 5; ++j) part[wave][j] = acc[j];
    __syncthreads();
    if (threadIdx.x == 0) {
      const int bid = blockIdx.y * gridDim.x + blockIdx.x;
      for (int j = 0; j < 5; ++j) {
        double t = 0.0;
        for (int w = 0; w < WAVES; ++w) t += part[w][j];
        red[5 * bid + j] = t;
      }
    }
  }
}

template <int ROWS, int WAVES, int COLS>
hipError_t launch_cheb_stream_shape(const float* xin, const float* rin, const float* din,
                                    const float* b, const uint8_t* mask, float* xout,
                                    float* rout, float* dout, int W, int H, int k, int seg,
                                    int first, float theta, const ChebCoef& cf, double* red,
                                    const float* dapp, const double* line, hipStream_t st) {
  constexpr int EX = 64 * COLS;
  const dim3 grid((W + EX - 2 * k - 1) / (EX - 2 * k), (H + seg - 1) / seg);
  if (red)
    k_cheb_stream<ROWS, WAVES, COLS, true><<<grid, WAVES * 64, 0, st>>>(
        xin, rin, din, b, mask, xout, rout, dout, W, H, k, seg, first, theta, cf, red, dapp,
        line);
  else
    k_cheb_stream<ROWS, WAVES, COLS, false><<<grid, WAVES * 64, 0, st>>>(
        xin, rin, din, b, mask, xout, rout, dout, W, H, k, seg, first, theta, cf, nullptr,
        dapp, line);
  return hipGetLastError();
}

// One Jacobi sweep over a wave's register rows, descending, in place:
// p' = ((((x1 + x2) + y1) + y2) - b) * 0.25 (shader.Poisson.comp:37) with the
// centre row one register row up; b moves with p.  With RESID, the FP64
// residual b - ((((x1 + x2) + y1) + y2) - 4 p) of the sweep's input iterate
// is accumulated over the fluid cells of the written interior.
template <bool RESID, int ROWS, int COLS>
__device__ __forceinline__ void jacobi_pass(float (&P)[ROWS][COLS], float (&B)[ROWS][COLS],
                                            int (&SL)[ROWS][COLS], const float (&p0)[COLS],
                                            const float (&p1)[COLS], const float (&p2)[COLS],
                                            const int (&ps)[COLS], int row_new0, int gx0,
                                            int tcol0, int k, int Y0, int Y1, int W, int EX,
                                            double& sr, double& srr, bool special) {
#pragma unroll
  for (int i = ROWS - 1; i >= 0; --i) {
    float vc[COLS], va[COLS], vb[COLS], bc[COLS];
    int sc[COLS], sa[COLS], sb[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      vc[c] = i >= 1 ? P[i - 1][c] : p1[c];
      va[c] = i >= 2 ? P[i - 2][c] : (i == 1 ? p1[c] : p0[c]);
      vb[c] = P[i][c];
      bc[c] = i >= 1 ? B[i - 1][c] : p2[c];
      sc[c] = i >= 1 ? SL[i - 1][c] : (ps[c] >> 1) & 1;
      sa[c] = i >= 2 ? SL[i - 2][c] : (i == 1 ? (ps[c] >> 1) & 1 : ps[c] & 1);
      sb[c] = SL[i][c];
    }
    const float lft = lane_from_left(vc[COLS - 1]);
    const float rgt = lane_from_right(vc[0]);
    int slf = 0, srg = 0;
    if (special) {
      slf = ilane_from_left(sc[COLS - 1]);
      srg = ilane_from_right(sc[0]);
    }
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const float v = vc[c];
      float x1 = c > 0 ? vc[c - 1] : lft;
      float x2 = c < COLS - 1 ? vc[c + 1] : rgt;
      float y1 = va[c];
      float y2 = vb[c];
      if (special) {
        const int s1 = c > 0 ? sc[c - 1] : slf;
        const int s2 = c < COLS - 1 ? sc[c + 1] : srg;
        x1 = s1 ? v : x1;
        x2 = s2 ? v : x2;
        y1 = sa[c] ? v : y1;
        y2 = sb[c] ? v : y2;
      }
      P[i][c] = ((((x1 + x2) + y1) + y2) - bc[c]) * 0.25f;
      if (RESID) {
        const int gy = row_new0 + i, gx = gx0 + c, tcol = tcol0 + c;
        if (gy >= Y0 && gy < Y1 && tcol >= k && tcol < EX - k && gx >= 0 && gx < W &&
            !sc[c]) {
          const double rr = (double)bc[c] - (((((double)x1 + (double)x2) + (double)y1) +
                                              (double)y2) -
                                             4.0 * (double)v);
          sr += rr;
          srr += rr * rr;
        }
      }
      B[i][c] = bc[c];
      SL[i][c] = sc[c];
    }
  }
}

// k Jacobi sweeps per launch, streaming; red != nullptr: the last sweep
// also writes the FP64 residual sums of its input iterate,
// red[2 blk + {sum r, sum r^2}] over the fluid cells the launch writes.
// done (nullable): a device flag; when non-zero the launch returns at once.
template <int ROWS, int WAVES, int COLS>
__global__ void __launch_bounds__(WAVES * 64)
    k_jacobi_stream(const float* __restrict__ pin, const float* __restrict__ b,
                    const uint8_t* __restrict__ mask, float* __restrict__ pout, int W, int H,
                    int k, int seg, double* __restrict__ red, const int* __restrict__ done) {
  constexpr int EX = 64 * COLS, S = ROWS * WAVES;
  static_assert(ROWS >= 2, "the pass reads two rows above a wave's first row");
  static_assert(WAVES >= 2, "wave WAVES - 1 keeps the carry in the registers wave 0 reads it into");
  if (done != nullptr && *done != 0) return;  // block-uniform early exit
  const int lane = threadIdx.x & 63;
  // the wave index is uniform across the wave: an SGPR, so branches on it are
  // scalar branches with no exec-mask merging
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tcol0 = COLS * lane;
  const int gx0 = (int)blockIdx.x * (EX - 2 * k) - k + tcol0;
  const int Y0 = (int)blockIdx.y * seg;
  const int Y1 = min(Y0 + seg, H);
  const int A = Y0 - k;
  const int nch = (Y1 - Y0 + 2 * k + S - 1) / S;
  const bool vec = COLS == 2 && (W % 2 == 0) && (k % 2 == 0);

  __shared__ Pub<EX, 3> pub[2][WAVES];
  __shared__ Pub<EX, 3> car[STREAM_KMAX];

  float P[ROWS][COLS], B[ROWS][COLS];
  int SL[ROWS][COLS];
  // the next chunk's inputs, loaded while the current chunk's sweeps run
  float Pn[ROWS][COLS], Bn[ROWS][COLS];
  int SLn[ROWS][COLS];
  double sr = 0.0, srr = 0.0;

  auto load_chunk = [&](int ch) {
    const int row0 = A + ch * S + wave * ROWS;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i;
      const bool row_in = gy >= 0 && gy < H;
      // unconditional loads from a clamped in-grid address, then selects
      if (vec) {
        const bool in = row_in && gx0 >= 0 && gx0 < W;
        const int idx = in ? gy * W + gx0 : 0;
        const float2 pv = *reinterpret_cast<const float2*>(pin + idx);
        const float2 bv = *reinterpret_cast<const float2*>(b + idx);
        const uint16_t mm = *reinterpret_cast<const uint16_t*>(mask + idx);
        Pn[i][0] = in ? pv.x : 0.0f;
        Pn[i][COLS - 1] = in ? pv.y : 0.0f;
        Bn[i][0] = in ? bv.x : 0.0f;
        Bn[i][COLS - 1] = in ? bv.y : 0.0f;
        SLn[i][0] = in ? (mm & 0xFF) != 0 : 1;
        SLn[i][COLS - 1] = in ? (mm >> 8) != 0 : 1;
      } else {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const int gx = gx0 + c;
          const bool in = row_in && gx >= 0 && gx < W;
          const int idx = in ? gy * W + gx : 0;
          const float pv = pin[idx], bv = b[idx];
          const uint8_t mv = mask[idx];
          Pn[i][c] = in ? pv : 0.0f;
          Bn[i][c] = in ? bv : 0.0f;
          SLn[i][c] = in ? (mv != 0) : 1;
        }
      }
    }
  };

  load_chunk(0);
  for (int ch = 0; ch < nch; ++ch) {
    const int row0 = A + ch * S + wave * ROWS;
#pragma unroll
    for (int i = 0; i < ROWS; ++i)
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        P[i][c] = Pn[i][c];
        B[i][c] = Bn[i][c];
        SL[i][c] = SLn[i][c];
      }
    if (ch + 1 < nch) load_chunk(ch + 1);

    for (int s = 0; s < k; ++s) {
      const int par = s & 1;
      float p0[COLS], p1[COLS], p2[COLS];
      int ps[COLS];
      if (wave == 0) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const bool have = ch > 0;
          p0[c] = have ? car[s].v[0][tcol0 + c] : 0.0f;
          p1[c] = have ? car[s].v[1][tcol0 + c] : 0.0f;
          p2[c] = have ? car[s].v[2][tcol0 + c] : 0.0f;
          ps[c] = have ? car[s].s[tcol0 + c] : 3;
        }
      }
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const uint8_t sb = (uint8_t)(SL[ROWS - 2][c] | (SL[ROWS - 1][c] << 1));
        pub[par][wave].v[0][tcol0 + c] = P[ROWS - 2][c];
        pub[par][wave].v[1][tcol0 + c] = P[ROWS - 1][c];
        pub[par][wave].v[2][tcol0 + c] = B[ROWS - 1][c];
        pub[par][wave].s[tcol0 + c] = sb;
        if (wave == WAVES - 1) {
          p0[c] = P[ROWS - 2][c];
          p1[c] = P[ROWS - 1][c];
          p2[c] = B[ROWS - 1][c];
          ps[c] = sb;
        }
      }
      __syncthreads();
      if (wave == WAVES - 1) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          car[s].v[0][tcol0 + c] = p0[c];
          car[s].v[1][tcol0 + c] = p1[c];
          car[s].v[2][tcol0 + c] = p2[c];
          car[s].s[tcol0 + c] = (uint8_t)ps[c];
        }
      }
      if (wave > 0) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          p0[c] = pub[par][wave - 1].v[0][tcol0 + c];
          p1[c] = pub[par][wave - 1].v[1][tcol0 + c];
          p2[c] = pub[par][wave - 1].v[2][tcol0 + c];
          ps[c] = pub[par][wave - 1].s[tcol0 + c];
        }
      }
      const int row_new0 = row0 - (s + 1);
      int anys = ps[0] | (COLS > 1 ? ps[COLS - 1] : 0);
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c) anys |= SL[i][c];
      const bool special = __any(anys != 0);
      if (red != nullptr && s == k - 1)
        jacobi_pass<true>(P, B, SL, p0, p1, p2, ps, row_new0, gx0, tcol0, k, Y0, Y1, W, EX,
                          sr, srr, special);
      else
        jacobi_pass<false>(P, B, SL, p0, p1, p2, ps, row_new0, gx0, tcol0, k, Y0, Y1, W, EX,
                           sr, srr, special);
    }

#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i - k;
      if (gy < Y0 || gy >= Y1) continue;
      if (vec) {
        if (tcol0 < k || tcol0 >= EX - k || gx0 < 0 || gx0 >= W) continue;
        *reinterpret_cast<float2*>(pout + gy * W + gx0) = make_float2(P[i][0], P[i][COLS - 1]);
        continue;
      }
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int tcol = tcol0 + c, gx = gx0 + c;
        if (tcol < k || tcol >= EX - k || gx < 0 || gx >= W) continue;
        pout[gy * W + gx] = P[i][c];
      }
    }
    __syncthreads();
  }

  if (red != nullptr) {
    __shared__ double part[WAVES][2];
    sr = jt::wave_sum(sr);
    srr = jt::wave_sum(srr);
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

template <int ROWS, int WAVES, int COLS>
hipError_t launch_jacobi_stream_shape(const float* pin, const float* b, const uint8_t* mask,
                                      float* pout, int W, int H, int k, int seg, double* red,
                                      const int* done, hipStream_t st) {
  constexpr int EX = 64 * COLS;
  const dim3 grid((W + EX - 2 * k - 1) / (EX - 2 * k), (H + seg - 1) / seg);
  k_jacobi_stream<ROWS, WAVES, COLS><<<grid, WAVES * 64, 0, st>>>(pin, b, mask, pout, W, H, k,
                                                                   seg, red, done);
  return hipGetLastError();
}

}  // namespace

// Streaming shapes (rows per wave, waves per block, columns per lane).
#define NX_STREAM_SHAPES(X) \
  X(4, 8, 2)                \
  X(2, 16, 2)               \
  X(4, 16, 2)               \
  X(8, 8, 2)

bool stream_cfg_supported(JacobiCfg c) {
  bool shape = false;
#define NX_STREAM_CHECK(R, WV, CL) shape |= (c.rows == R && c.waves == WV && c.cols == CL);
  NX_STREAM_SHAPES(NX_STREAM_CHECK)
#undef NX_STREAM_CHECK
  return shape && c.k >= 1 && c.k <= STREAM_KMAX && 64 * c.cols - 2 * c.k >= 8 && c.seg >= 1;
}

int stream_launch_blocks(int W, int H, JacobiCfg c) {
  const int VX = 64 * c.cols - 2 * c.k;
  return ((W + VX - 1) / VX) * ((H + c.seg - 1) / c.seg);
}

hipError_t launch_cheb_stream(hipStream_t st, JacobiCfg c, const float* xin, const float* rin,
                              const float* din, const float* b, const uint8_t* mask,
                              float* xout, float* rout, float* dout, int W, int H, int setup,
                              float theta, const ChebCoef& cf, double* red, const float* dapp,
                              const double* line) {
  if (!stream_cfg_supported(c)) return hipErrorInvalidValue;
#define NX_STREAM_LAUNCH(R, WV, CL)                                                          \
  if (c.rows == R && c.waves == WV && c.cols == CL)                                         \
    return launch_cheb_stream_shape<R, WV, CL>(xin, rin, din, b, mask, xout, rout, dout, W, \
                                               H, c.k, c.seg, setup, theta, cf, red, dapp,  \
                                               line, st);
  NX_STREAM_SHAPES(NX_STREAM_LAUNCH)
#undef NX_STREAM_LAUNCH
  return hipErrorInvalidValue;
}

hipError_t launch_jacobi_stream(const float* pin, const float* b, const uint8_t* mask,
                                float* pout, int W, int H, JacobiCfg c, double* red,
                                const int* done, hipStream_t st) {
  if (!stream_cfg_supported(c)) return hipErrorInvalidValue;
#define NX_STREAM_JLAUNCH(R, WV, CL)                                                        \
  if (c.rows == R && c.waves == WV && c.cols == CL)                                        \
    return launch_jacobi_stream_shape<R, WV, CL>(pin, b, mask, pout, W, H, c.k, c.seg, red, \
                                                 done, st);
  NX_STREAM_SHAPES(NX_STREAM_JLAUNCH)
#undef NX_STREAM_JLAUNCH
  return hipErrorInvalidValue;
}

}  // namespace nx
