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

// Solid flags of a wave's register rows are packed one bit per row per
// column: bit i of SLB[c] is set when the cell of register row i in column
// c is solid or outside the grid.  For a pass, E[c] = (SLB[c] << 2) | ps[c]
// extends that by the two rows above the wave's strip (bit 0: row -2,
// bit 1: row -1, from the previous wave or the carry), so bit j of E is row
// j - 2; after the pass the rows have moved down by one and SLB[c] becomes
// (E[c] >> 1) masked to ROWS bits.
template <int ROWS, int COLS>
__device__ __forceinline__ bool extend_flags(int (&E)[COLS], const int (&SLB)[COLS],
                                             const int (&ps)[COLS]) {
  int any = 0;
#pragma unroll
  for (int c = 0; c < COLS; ++c) {
    E[c] = (SLB[c] << 2) | ps[c];
    any |= E[c];
  }
  return any != 0;
}

template <int ROWS, int COLS>
__device__ __forceinline__ void shift_flags(int (&SLB)[COLS], const int (&E)[COLS]) {
#pragma unroll
  for (int c = 0; c < COLS; ++c) SLB[c] = (E[c] >> 1) & ((1 << ROWS) - 1);
}

// One Chebyshev pass over a wave's register rows, descending, in place.
// V is the stencil operand (d for iterations, x for setup and residual),
// (A, B) the other two arrays ((x, r) for iterations, (r, d) otherwise).
// p0..p3: the previous wave's (or the carry's) last rows; E: the extended
// solid flags, el / er those of the lanes to the left and right.  SP: with
// neighbour replacement (some flag is set in the wave); without it, no
// select is evaluated and the lane shifts fold into the first addition.
template <int KIND, bool SP, int ROWS, int COLS>
__device__ __forceinline__ void cheb_pass(float (&X)[ROWS][COLS], float (&R)[ROWS][COLS],
                                          float (&D)[ROWS][COLS], const int (&E)[COLS], int el,
                                          int er, const float (&p0)[COLS],
                                          const float (&p1)[COLS], const float (&p2)[COLS],
                                          const float (&p3)[COLS], float theta, float cdd,
                                          float crr, int row_new0, int gx0, int tcol0, int k,
                                          int Y0, int Y1, int W, int EX,
                                          const float (&BR)[ROWS][COLS], double (&acc)[5]) {
  float(&V)[ROWS][COLS] = KIND == P_ITER ? D : X;
  float(&Aa)[ROWS][COLS] = KIND == P_ITER ? X : R;
  float(&Bb)[ROWS][COLS] = KIND == P_ITER ? R : D;
#pragma unroll
  for (int i = ROWS - 1; i >= 0; --i) {
    float vc[COLS], va[COLS], vb[COLS], ac[COLS], bc[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      vc[c] = i >= 1 ? V[i - 1][c] : p1[c];
      va[c] = i >= 2 ? V[i - 2][c] : (i == 1 ? p1[c] : p0[c]);
      vb[c] = V[i][c];
      ac[c] = i >= 1 ? Aa[i - 1][c] : p2[c];
      bc[c] = i >= 1 ? Bb[i - 1][c] : p3[c];
    }
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const float v = vc[c];
      // the strip's outermost lanes read 0 across the wave edge: columns
      // outside the ones the launch keeps
      float x1 = c > 0 ? vc[c - 1] : jt::lane_shr0(vc[COLS - 1]);
      float x2 = c < COLS - 1 ? vc[c + 1] : jt::lane_shl0(vc[0]);
      float y1 = va[c];
      float y2 = vb[c];
      // register row j is bit j + 2 of E: the centre (row i - 1) is bit
      // i + 1, y1 (row i - 2) bit i, y2 (row i) bit i + 2
      const int sc = SP ? (E[c] >> (i + 1)) & 1 : 0;
      if (SP) {
        const int s1 = ((c > 0 ? E[c - 1] : el) >> (i + 1)) & 1;
        const int s2 = ((c < COLS - 1 ? E[c + 1] : er) >> (i + 1)) & 1;
        x1 = s1 ? v : x1;
        x2 = s2 ? v : x2;
        y1 = ((E[c] >> i) & 1) ? v : y1;
        y2 = ((E[c] >> (i + 2)) & 1) ? v : y2;
      }
      if (KIND == P_SETUP) {
        // 4 v is exact: the fused multiply-add rounds once, like the
        // subtraction of the product
        const float lv = __builtin_fmaf(-4.0f, v, ((y1 + x1) + x2) + y2);
        const float rn = lv - ac[c];  // r holds b before the setup
        X[i][c] = v;
        R[i][c] = rn;
        D[i][c] = rn / theta;
      } else if (KIND == P_ITER) {
        const float lv = __builtin_fmaf(-4.0f, v, ((y1 + x1) + x2) + y2);
        const float rn = bc[c] + lv;
        X[i][c] = ac[c] + v;
        R[i][c] = rn;
        D[i][c] = __builtin_fmaf(cdd, v, crr * rn);
      } else {
        const int gy = row_new0 + i, gx = gx0 + c, tcol = tcol0 + c;
        if (gy >= Y0 && gy < Y1 && tcol >= k && tcol < EX - k && gx >= 0 && gx < W && !sc) {
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
    }
  }
}

// VEC: float2 accesses (COLS == 2, even width and k: every lane's column
// pair is aligned and wholly inside or outside the grid and the columns
// the launch keeps).
template <int ROWS, int WAVES, int COLS, bool RES, bool VEC>
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
  static_assert(!VEC || COLS == 2, "float2 accesses take two columns per lane");
  const float a_line = dapp != nullptr ? (float)line[0] : 0.0f;

  __shared__ Pub<EX> pub[2][WAVES];
  // the per-pass carry, k entries in dynamic LDS (sized by the launch)
  extern __shared__ __align__(16) unsigned char nx_stream_carry[];
  Pub<EX>* car = reinterpret_cast<Pub<EX>*>(nx_stream_carry);

  float X[ROWS][COLS], R[ROWS][COLS], D[ROWS][COLS];
  int SLB[COLS];  // packed solid flags, bit i: register row i
  float BR[ROWS][COLS];  // b of the rows the residual pass writes (RES)
  // The next chunk's raw inputs, loaded from clamped in-grid addresses
  // while the current chunk's passes run: nothing reads them before the
  // next chunk starts, so the loads stay in flight across the passes'
  // barriers.  Cells outside the grid are replaced when the chunk starts.
  // With RES also b of the rows the last pass writes (register row i holds
  // grid row - k after the k passes).
  float Xr[ROWS][COLS], Rr[ROWS][COLS], Dr[ROWS][COLS];
  float Br[RES ? ROWS : 1][RES ? COLS : 1];
  uint32_t Mr[ROWS][VEC ? 1 : COLS];  // VEC: both mask bytes of the pair
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const float* rsrc = first ? b : rin;  // the first launch holds b in r

  auto load_chunk = [&](int ch) {
    const int row0 = A + ch * S + wave * ROWS;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int cy = min(max(row0 + i, 0), H - 1);
      const int cyr = min(max(row0 + i - k, 0), H - 1);
      if (VEC) {
        const int idx = cy * W + min(max(gx0, 0), W - 2);
        const float2 xv = *reinterpret_cast<const float2*>(xin + idx);
        const float2 rv = *reinterpret_cast<const float2*>(rsrc + idx);
        const float2 dv = *reinterpret_cast<const float2*>(din + idx);
        Mr[i][0] = *reinterpret_cast<const uint16_t*>(mask + idx);
        Xr[i][0] = xv.x;
        Xr[i][COLS - 1] = xv.y;
        Rr[i][0] = rv.x;
        Rr[i][COLS - 1] = rv.y;
        Dr[i][0] = dv.x;
        Dr[i][COLS - 1] = dv.y;
        if (RES) {
          const float2 bv =
              *reinterpret_cast<const float2*>(b + cyr * W + min(max(gx0, 0), W - 2));
          Br[RES ? i : 0][0] = bv.x;
          Br[RES ? i : 0][RES ? COLS - 1 : 0] = bv.y;
        }
      } else {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const int cx = min(max(gx0 + c, 0), W - 1);
          Xr[i][c] = xin[cy * W + cx];
          Rr[i][c] = rsrc[cy * W + cx];
          Dr[i][c] = din[cy * W + cx];
          Mr[i][VEC ? 0 : c] = mask[cy * W + cx];
          if (RES) Br[RES ? i : 0][RES ? c : 0] = b[cyr * W + cx];
        }
      }
    }
  };

  load_chunk(0);
  for (int ch = 0; ch < nch; ++ch) {
    const int row0 = A + ch * S + wave * ROWS;  // level-0 row of register row 0
    // the chunk's registers from its raw inputs
#pragma unroll
    for (int c = 0; c < COLS; ++c) SLB[c] = 0;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i, gyr = gy - k;
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gx = gx0 + c;
        const bool col_in = gx >= 0 && gx < W;
        const bool in = col_in && gy >= 0 && gy < H;
        const uint32_t m = VEC ? (c == 0 ? Mr[i][0] & 0xFF : Mr[i][0] >> 8) : Mr[i][VEC ? 0 : c];
        X[i][c] = in ? Xr[i][c] : 0.0f;
        R[i][c] = in ? Rr[i][c] : 0.0f;
        D[i][c] = (in && !first) ? Dr[i][c] : 0.0f;
        SLB[c] |= (in ? m != 0 : 1) << i;
        BR[i][c] = (RES && col_in && gyr >= 0 && gyr < H) ? Br[RES ? i : 0][RES ? c : 0] : 0.0f;
      }
    }
    // a pending line-search step x += a dapp on the fluid cells (the first
    // launch of a run); its loads are issued together, then applied
    if (a_line != 0.0f) {
      float av[ROWS][COLS];
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c)
          av[i][c] = dapp[min(max(row0 + i, 0), H - 1) * W + min(max(gx0 + c, 0), W - 1)];
#pragma unroll
      for (int i = 0; i < ROWS; ++i)
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const int gy = row0 + i, gx = gx0 + c;
          if (gy >= 0 && gy < H && gx >= 0 && gx < W && !((SLB[c] >> i) & 1))
            X[i][c] = X[i][c] + a_line * av[i][c];
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
        const uint8_t sb = (uint8_t)((SLB[c] >> (ROWS - 2)) & 3);
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
      int E[COLS];
      const bool special = __any(extend_flags<ROWS, COLS>(E, SLB, ps));
      const float cdd = it ? cf.cd[s] : 0.0f, crr = it ? cf.cr[s] : 0.0f;
      if (special) {
        const int el = jt::ilane_from_left(E[COLS - 1]), er = jt::ilane_from_right(E[0]);
        cheb_pass<KIND, true>(X, R, D, E, el, er, p0, p1, p2, p3, theta, cdd, crr,
                              row0 - (s + 1), gx0, tcol0, k, Y0, Y1, W, EX, BR, acc);
      } else {
        cheb_pass<KIND, false>(X, R, D, E, 0, 0, p0, p1, p2, p3, theta, cdd, crr,
                               row0 - (s + 1), gx0, tcol0, k, Y0, Y1, W, EX, BR, acc);
      }
      shift_flags<ROWS, COLS>(SLB, E);
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
      if (VEC) {
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
  const bool vec = COLS == 2 && W % 2 == 0 && k % 2 == 0;
  const size_t carry = (size_t)k * sizeof(Pub<EX>);
#define NX_CHEB_STREAM_GO(RESV, VECV)                                                   \
  k_cheb_stream<ROWS, WAVES, COLS, RESV, VECV><<<grid, WAVES * 64, carry, st>>>(        \
      xin, rin, din, b, mask, xout, rout, dout, W, H, k, seg, first, theta, cf, red, dapp, \
      line)
  if (red && vec)
    NX_CHEB_STREAM_GO(true, COLS == 2);
  else if (red)
    NX_CHEB_STREAM_GO(true, false);
  else if (vec)
    NX_CHEB_STREAM_GO(false, COLS == 2);
  else
    NX_CHEB_STREAM_GO(false, false);
#undef NX_CHEB_STREAM_GO
  return hipGetLastError();
}

// One Jacobi sweep over a wave's register rows, descending, in place:
// p' = ((((x1 + x2) + y1) + y2) - b) * 0.25 (shader.Poisson.comp:37) with the
// centre row one register row up; b moves with p.  With RESID, the FP64
// residual b - ((((x1 + x2) + y1) + y2) - 4 p) of the sweep's input iterate
// is accumulated over the fluid cells of the written interior.  E, el, er
// and SP as for cheb_pass.
template <bool RESID, bool SP, int ROWS, int COLS>
__device__ __forceinline__ void jacobi_pass(float (&P)[ROWS][COLS], float (&B)[ROWS][COLS],
                                            const int (&E)[COLS], int el, int er,
                                            const float (&p0)[COLS], const float (&p1)[COLS],
                                            const float (&p2)[COLS], int row_new0, int gx0,
                                            int tcol0, int k, int Y0, int Y1, int W, int EX,
                                            double& sr, double& srr) {
#pragma unroll
  for (int i = ROWS - 1; i >= 0; --i) {
    float vc[COLS], va[COLS], vb[COLS], bc[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      vc[c] = i >= 1 ? P[i - 1][c] : p1[c];
      va[c] = i >= 2 ? P[i - 2][c] : (i == 1 ? p1[c] : p0[c]);
      vb[c] = P[i][c];
      bc[c] = i >= 1 ? B[i - 1][c] : p2[c];
    }
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
      const float v = vc[c];
      float x1 = c > 0 ? vc[c - 1] : jt::lane_shr0(vc[COLS - 1]);
      float x2 = c < COLS - 1 ? vc[c + 1] : jt::lane_shl0(vc[0]);
      float y1 = va[c];
      float y2 = vb[c];
      const int sc = SP ? (E[c] >> (i + 1)) & 1 : 0;
      if (SP) {
        const int s1 = ((c > 0 ? E[c - 1] : el) >> (i + 1)) & 1;
        const int s2 = ((c < COLS - 1 ? E[c + 1] : er) >> (i + 1)) & 1;
        x1 = s1 ? v : x1;
        x2 = s2 ? v : x2;
        y1 = ((E[c] >> i) & 1) ? v : y1;
        y2 = ((E[c] >> (i + 2)) & 1) ? v : y2;
      }
      P[i][c] = ((((x1 + x2) + y1) + y2) - bc[c]) * 0.25f;
      if (RESID) {
        const int gy = row_new0 + i, gx = gx0 + c, tcol = tcol0 + c;
        if (gy >= Y0 && gy < Y1 && tcol >= k && tcol < EX - k && gx >= 0 && gx < W && !sc) {
          const double rr = (double)bc[c] - (((((double)x1 + (double)x2) + (double)y1) +
                                              (double)y2) -
                                             4.0 * (double)v);
          sr += rr;
          srr += rr * rr;
        }
      }
      B[i][c] = bc[c];
    }
  }
}

// k Jacobi sweeps per launch, streaming; red != nullptr: the last sweep
// also writes the FP64 residual sums of its input iterate,
// red[2 blk + {sum r, sum r^2}] over the fluid cells the launch writes.
// done (nullable): a device flag; when non-zero the launch returns at once.
template <int ROWS, int WAVES, int COLS, bool VEC>
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
  static_assert(!VEC || COLS == 2, "float2 accesses take two columns per lane");

  __shared__ Pub<EX, 3> pub[2][WAVES];
  extern __shared__ __align__(16) unsigned char nx_stream_carry[];
  Pub<EX, 3>* car = reinterpret_cast<Pub<EX, 3>*>(nx_stream_carry);

  float P[ROWS][COLS], B[ROWS][COLS];
  int SLB[COLS];  // packed solid flags, bit i: register row i
  // the next chunk's raw inputs (clamped in-grid addresses), loaded while
  // the current chunk's sweeps run and taken when the next chunk starts
  float Pr[ROWS][COLS], Br[ROWS][COLS];
  uint32_t Mr[ROWS][VEC ? 1 : COLS];  // VEC: both mask bytes of the pair
  double sr = 0.0, srr = 0.0;

  auto load_chunk = [&](int ch) {
    const int row0 = A + ch * S + wave * ROWS;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int cy = min(max(row0 + i, 0), H - 1);
      if (VEC) {
        const int idx = cy * W + min(max(gx0, 0), W - 2);
        const float2 pv = *reinterpret_cast<const float2*>(pin + idx);
        const float2 bv = *reinterpret_cast<const float2*>(b + idx);
        Mr[i][0] = *reinterpret_cast<const uint16_t*>(mask + idx);
        Pr[i][0] = pv.x;
        Pr[i][COLS - 1] = pv.y;
        Br[i][0] = bv.x;
        Br[i][COLS - 1] = bv.y;
      } else {
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
          const int idx = cy * W + min(max(gx0 + c, 0), W - 1);
          Pr[i][c] = pin[idx];
          Br[i][c] = b[idx];
          Mr[i][VEC ? 0 : c] = mask[idx];
        }
      }
    }
  };

  load_chunk(0);
  for (int ch = 0; ch < nch; ++ch) {
    const int row0 = A + ch * S + wave * ROWS;
#pragma unroll
    for (int c = 0; c < COLS; ++c) SLB[c] = 0;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i;
#pragma unroll
      for (int c = 0; c < COLS; ++c) {
        const int gx = gx0 + c;
        const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
        const uint32_t m = VEC ? (c == 0 ? Mr[i][0] & 0xFF : Mr[i][0] >> 8) : Mr[i][VEC ? 0 : c];
        P[i][c] = in ? Pr[i][c] : 0.0f;
        B[i][c] = in ? Br[i][c] : 0.0f;
        SLB[c] |= (in ? m != 0 : 1) << i;
      }
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
        const uint8_t sb = (uint8_t)((SLB[c] >> (ROWS - 2)) & 3);
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
      int E[COLS];
      const bool special = __any(extend_flags<ROWS, COLS>(E, SLB, ps));
      if (red != nullptr && s == k - 1) {
        const int el = jt::ilane_from_left(E[COLS - 1]), er = jt::ilane_from_right(E[0]);
        jacobi_pass<true, true>(P, B, E, el, er, p0, p1, p2, row_new0, gx0, tcol0, k, Y0, Y1,
                                W, EX, sr, srr);
      } else if (special) {
        const int el = jt::ilane_from_left(E[COLS - 1]), er = jt::ilane_from_right(E[0]);
        jacobi_pass<false, true>(P, B, E, el, er, p0, p1, p2, row_new0, gx0, tcol0, k, Y0, Y1,
                                 W, EX, sr, srr);
      } else {
        jacobi_pass<false, false>(P, B, E, 0, 0, p0, p1, p2, row_new0, gx0, tcol0, k, Y0, Y1,
                                  W, EX, sr, srr);
      }
      shift_flags<ROWS, COLS>(SLB, E);
    }

#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
      const int gy = row0 + i - k;
      if (gy < Y0 || gy >= Y1) continue;
      if (VEC) {
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
  const size_t carry = (size_t)k * sizeof(Pub<EX, 3>);
  if (COLS == 2 && W % 2 == 0 && k % 2 == 0)
    k_jacobi_stream<ROWS, WAVES, COLS, COLS == 2><<<grid, WAVES * 64, carry, st>>>(
        pin, b, mask, pout, W, H, k, seg, red, done);
  else
    k_jacobi_stream<ROWS, WAVES, COLS, false><<<grid, WAVES * 64, carry, st>>>(
        pin, b, mask, pout, W, H, k, seg, red, done);
  return hipGetLastError();
}

}  // namespace

// Streaming shapes (rows per wave, waves per block, columns per lane).
#define NX_STREAM_SHAPES(X) \
  X(4, 8, 2)                \
  X(2, 16, 2)               \
  X(4, 16, 2)               \
  X(8, 8, 2)                \
  X(8, 4, 2)

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
