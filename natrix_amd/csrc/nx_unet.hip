// U-Net building blocks for the learned pressure guess (gfx950).
//
// Activations are NHWC bfloat16, 3x3 conv weights KRSC bfloat16
// (w[cout][tap][cin], tap = 3*ky + kx), accumulation float32 on the matrix
// cores with v_mfma_f32_32x32x16_bf16:
//   * forward: implicit GEMM, M = pixels, N = output channels,
//     K = 9 * input channels.  A lane's A fragment (8 consecutive K = 8
//     channels of one tap of one pixel) and B fragment (8 input channels of
//     one output channel) are each one contiguous 16-byte load, so the
//     operands stream from the caches straight into registers.  Bias and
//     ReLU are fused into the epilogue.
//   * input gradient: the same kernel on the output gradient with the
//     weights flipped and transposed (w'[cin][8-tap][cout]).
//   * weight gradient: GEMM over the pixel dimension (K = pixels); 64-pixel
//     chunks of the output gradient and the tap-shifted input are staged in
//     LDS transposed to [channel][pixel] so each lane again reads one
//     16-byte fragment; the pixel range is split across blocks and the
//     float32 partials are summed in a fixed order.
// The MFMA operand layout used (32x32x16): lane l holds row (l mod 32) of A
// and column (l mod 32) of B, with K elements 8*(l/32) .. 8*(l/32)+7; the
// accumulator's 16 registers of lane l hold column (l mod 32), rows
// 8*(r/4) + 4*(l/32) + (r mod 4).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/natrix.h"
#include "nx_common.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ bf16x8 ld_frag(const uint16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ bf16x8 zero_frag() {
  return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
}

__device__ __forceinline__ float bf(uint16_t h) { return nx::bf16_to_f32(h); }
__device__ __forceinline__ uint16_t tobf(float f) { return (uint16_t)nx::f32_to_bf16(f); }

// ---------------------------------------------------------------- forward
// WT (Cin % 16 == 0 only): the weights are tiled as wt[tap][ci / 16][co][16]
// (k_tile_weights), so the B fragments of one K step, 8 channels of 32
// output channels per half-wave, are one contiguous 1 KiB block instead of
// 16-byte pieces of 32 rows 9 Cin apart; the arithmetic is the same.
// grid: (ceil(M / (4*32*MT)), Cout / (32*NT), splits); block 256 = 4 waves
// along M.  With part == nullptr (one split) the block runs the whole K loop
// and writes bias + ReLU in bf16; otherwise (Cin % 16 == 0) split z runs
// K steps [z * steps_per_split, ...) of 16 channels of one tap each and
// writes its float32 partial sums to part[z][m][co] for k_conv_finish.
// (Letting the last split of each tile sum the tile's partials instead, with
// an arrival counter, measured several times slower: one workgroup per tile
// then reads every split's partial of the tile, where the finish pass
// spreads those reads over the whole chip.)
// CAT (decoder convolutions, Cin % 16 == 0): the input is the channel
// concatenation of the 2x nearest upsampling of x (C1 channels, an
// (H / 2) x (W / 2) image) and x2 (Cin - C1 channels, H x W), read where
// the fragments are loaded instead of from a materialised copy; a K step's
// 16 channels lie in one of the two (C1 % 16 == 0), so every value read is
// the one the upsampled, concatenated tensor would hold.
template <int MT, int NT, bool WT, bool CAT = false>
__global__ void __launch_bounds__(256)
    k_conv3x3(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
              const float* __restrict__ bias, uint16_t* __restrict__ y, int N,
              int H, int W, int Cin, int Cout, int relu, float* __restrict__ part,
              int steps_per_split, const uint16_t* __restrict__ x2 = nullptr,
              int C1 = 0) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, kb = lane >> 5;
  const long long M = (long long)N * H * W;
  const long long m_base = ((long long)blockIdx.x * 4 + wave) * 32 * MT;
  const int n_base = blockIdx.y * 32 * NT;
  int py[MT], px[MT];
  long long pimg[MT];   // first pixel of the pixel's image
  long long pimgh[MT];  // the same in the half-resolution image (CAT)
  bool valid[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) {
    const long long m = m_base + 32 * i + r;
    valid[i] = m < M;
    const long long mm = valid[i] ? m : 0;
    const int hw = H * W;
    const long long img = mm / hw;
    const int rem = (int)(mm - img * hw);
    py[i] = rem / W;
    px[i] = rem - py[i] * W;
    pimg[i] = img * hw;
    pimgh[i] = img * (long long)(H >> 1) * (W >> 1);
  }
  f32x16 acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  const int K = 9 * Cin;
  if ((Cin & 15) == 0) {
    // K steps of 16 channels of one tap, in order, unrolled by two so the
    // fragments of the next step load while the MFMAs of the current one
    // run.  Loads are unconditional: an out-of-image pixel reads x[0] and
    // its fragment is zeroed by a mask; the prefetch past the last step
    // re-reads that step (clamped) and is never used.
    const int cps = Cin >> 4;  // K steps per tap
    const int ks0 = part ? blockIdx.z * steps_per_split : 0;
    const int ks1 = part ? min(ks0 + steps_per_split, 9 * cps) : 9 * cps;
    uint4 aA[MT], bA[NT], aB[MT], bB[NT];
    uint32_t mA[MT], mB[MT];
#define NX_CONV_LOAD(A_, B_, M_, KS_)                                                  \
  {                                                                                    \
    const int ks_ = (KS_);                                                             \
    const int tap_ = ks_ / cps;                                                        \
    const int ci_ = (ks_ - tap_ * cps) * 16 + kb * 8;                                  \
    const int dy_ = tap_ / 3 - 1, dx_ = tap_ % 3 - 1;                                  \
    const bool up_ = CAT && ci_ < C1;                                                  \
    _Pragma("unroll") for (int i = 0; i < MT; ++i) {                                   \
      const int yy = py[i] + dy_, xx = px[i] + dx_;                                    \
      const bool ok = valid[i] && yy >= 0 && yy < H && xx >= 0 && xx < W;              \
      long long off;                                                                   \
      if (!CAT)                                                                        \
        off = ok ? (pimg[i] + (long long)yy * W + xx) * Cin + ci_ : 0;                 \
      else if (up_)                                                                    \
        off = ok ? (pimgh[i] + (long long)(yy >> 1) * (W >> 1) + (xx >> 1)) * C1 + ci_ : 0; \
      else                                                                             \
        off = ok ? (pimg[i] + (long long)yy * W + xx) * (Cin - C1) + (ci_ - C1) : 0;   \
      A_[i] = *reinterpret_cast<const uint4*>((CAT && !up_ ? x2 : x) + off);           \
      M_[i] = ok ? 0xFFFFFFFFu : 0u;                                                   \
    }                                                                                  \
    const int s_ = ks_ - tap_ * cps;                                                   \
    _Pragma("unroll") for (int j = 0; j < NT; ++j) B_[j] = *reinterpret_cast<const uint4*>( \
        WT ? w + ((long long)(tap_ * cps + s_) * Cout + n_base + 32 * j + r) * 16 + kb * 8 \
           : w + ((long long)(n_base + 32 * j + r) * 9 + tap_) * Cin + ci_);            \
  }
#define NX_CONV_MMA(A_, B_, M_)                                                        \
  {                                                                                    \
    _Pragma("unroll") for (int i = 0; i < MT; ++i) {                                   \
      const bf16x8 af = __builtin_bit_cast(                                            \
          bf16x8, make_uint4(A_[i].x & M_[i], A_[i].y & M_[i], A_[i].z & M_[i],        \
                             A_[i].w & M_[i]));                                        \
      _Pragma("unroll") for (int j = 0; j < NT; ++j) acc[i][j] =                       \
          __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, __builtin_bit_cast(bf16x8, B_[j]), \
                                                  acc[i][j], 0, 0, 0);                 \
    }                                                                                  \
  }
    int ks = ks0;
    if (ks < ks1) {
      NX_CONV_LOAD(aA, bA, mA, ks);
      for (; ks + 1 < ks1; ks += 2) {
        NX_CONV_LOAD(aB, bB, mB, ks + 1);
        NX_CONV_MMA(aA, bA, mA);
        NX_CONV_LOAD(aA, bA, mA, min(ks + 2, ks1 - 1));
        NX_CONV_MMA(aB, bB, mB);
      }
      if (ks < ks1) NX_CONV_MMA(aA, bA, mA);
    }
#undef NX_CONV_LOAD
#undef NX_CONV_MMA
  } else {
    // Cin a multiple of 8: one 8-wide fragment never crosses a tap
    for (int k0 = 0; k0 < K; k0 += 16) {
      const int kk = k0 + kb * 8;
      const int tap = kk / Cin, ci = kk - tap * Cin;
      bf16x8 a[MT], b[NT];
#pragma unroll
      for (int i = 0; i < MT; ++i) {
        a[i] = zero_frag();
        if (tap < 9) {
          const int yy = py[i] + tap / 3 - 1, xx = px[i] + tap % 3 - 1;
          if (valid[i] && yy >= 0 && yy < H && xx >= 0 && xx < W)
            a[i] = ld_frag(x + (pimg[i] + (long long)yy * W + xx) * Cin + ci);
        }
      }
#pragma unroll
      for (int j = 0; j < NT; ++j)
        b[j] = tap < 9 ? ld_frag(w + ((long long)(n_base + 32 * j + r) * 9 + tap) * Cin + ci)
                       : zero_frag();
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }

  if (part) {
    float* pz = part + (long long)blockIdx.z * M * Cout;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int co = n_base + 32 * j + r;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const long long m = m_base + 32 * i + 8 * (e >> 2) + 4 * kb + (e & 3);
          if (m < M) pz[m * Cout + co] = acc[i][j][e];
        }
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int co = n_base + 32 * j + r;
      const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const long long m = m_base + 32 * i + 8 * (e >> 2) + 4 * kb + (e & 3);
        if (m < M) {
          float v = acc[i][j][e] + bv;
          if (relu) v = fmaxf(v, 0.0f);
          y[m * Cout + co] = tobf(v);
        }
      }
    }
}

// y[m][co] = bf16(relu(sum over splits z in order of part[z][m][co] + bias))
__global__ void __launch_bounds__(256)
    k_conv_finish(const float* __restrict__ part, const float* __restrict__ bias,
                  uint16_t* __restrict__ y, long long MC, int Cout, int splits, int relu) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= MC) return;
  float v = 0.0f;
  for (int z = 0; z < splits; ++z) v += part[z * MC + i];
  v += bias ? bias[i % Cout] : 0.0f;
  if (relu) v = fmaxf(v, 0.0f);
  y[i] = tobf(v);
}

// In-workgroup split-K (Cin % 16 == 0): a workgroup of S = used waves
// computes one 32 x 32 output tile (32 pixels, 32 output channels); wave z
// runs K steps [z per, min(z per + per, steps)) of 16 channels of one tap,
// exactly the steps split z of k_conv3x3 runs, in phases of KS_MAX steps
// (at most KS_PHASES) whose fragment loads are all issued before the
// phase's first MFMA.  The S partial
// tiles are then summed in LDS in split order from 0, the bias added, ReLU
// applied and the tile stored in bf16: the operations of the split kernel's
// partials summed by k_conv_finish, in the same order (one MFMA computes
// each output element's 16-term dot product the same way whatever else is
// in its tile), so the result is bitwise the same, with no float32 partial
// leaving the chip and no second launch.  Dynamic LDS: S x 4 KiB.
constexpr int KS_MAX = 9, KS_PHASES = 2, KS_WAVES = 16;

// POOL (W % 16 == 0, H % 2 == 0): the tile's 32 pixels are 16 columns of 2
// rows instead of 32 consecutive pixels, and the epilogue also writes the
// 2 x 2 average pool of its bf16 outputs to yp, summed ((x00 + x01) + x10)
// + x11 and scaled by 0.25 as k_avgpool2 does.
template <bool WT, bool CAT, int PH, bool POOL>
__global__ void __launch_bounds__(64 * KS_WAVES)
    k_conv3x3_ks(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                 const float* __restrict__ bias, uint16_t* __restrict__ y, int N, int H,
                 int W, int Cin, int Cout, int relu, int per, int steps,
                 const uint16_t* __restrict__ x2, int C1, uint16_t* __restrict__ yp) {
  extern __shared__ float ks_part[];  // [S][32 rows][32 columns] (+ POOL: [32][32])
  const int lane = threadIdx.x & 63;
  const int z = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int S = blockDim.x >> 6;
  const int r = lane & 31, kb = lane >> 5;
  const long long M = (long long)N * H * W;
  const long long m_base = (long long)blockIdx.x * 32;
  const int n_base = blockIdx.y * 32;
  const int hw = H * W;
  // POOL tile: image ti, rows 2 ty, 2 ty + 1, columns 16 tx .. 16 tx + 15
  const int tw = W >> 4;
  const long long ti = POOL ? (long long)blockIdx.x / (tw * (H >> 1)) : 0;
  const int trem = POOL ? (int)(blockIdx.x - ti * (tw * (H >> 1))) : 0;
  const int ty = POOL ? trem / tw : 0, tx = POOL ? trem - ty * tw : 0;
  auto pixel_of = [&](int row) -> long long {  // tile row (MFMA row) -> pixel index
    return POOL ? ti * hw + (long long)(2 * ty + (row >> 4)) * W + 16 * tx + (row & 15)
                : m_base + row;
  };
  const long long m = pixel_of(r);
  const bool valid = m < M;
  const long long mm = valid ? m : 0;
  const long long img = mm / hw;
  const int rem = (int)(mm - img * hw);
  const int py = rem / W, px = rem - py * W;
  const long long pimg = img * hw, pimgh = img * (long long)(H >> 1) * (W >> 1);
  const int cps = Cin >> 4;
  const int ks0 = z * per, nst = min(per, steps - ks0);  // >= 1: z < used
  f32x16 acc;
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
  // the steps in PH phases of at most KS_MAX: all loads of a phase, then
  // its MFMAs in step order (the accumulation order of the split kernel)
#pragma unroll
  for (int q = 0; q < PH; ++q) {
    const int ph = q * KS_MAX;
    if (ph >= nst) continue;  // wave-uniform
    const int np = min(KS_MAX, nst - ph);
    uint4 a[KS_MAX], b[KS_MAX];
    uint32_t msk[KS_MAX];
#pragma unroll
    for (int t = 0; t < KS_MAX; ++t) {
      if (t >= np) continue;  // wave-uniform; every index stays static
      const int ks = ks0 + ph + t;
      const int tap = ks / cps, sblk = ks - tap * cps;
      const int ci = sblk * 16 + kb * 8;
      const int yy = py + tap / 3 - 1, xx = px + tap % 3 - 1;
      const bool ok = valid && yy >= 0 && yy < H && xx >= 0 && xx < W;
      const bool up = CAT && ci < C1;
      long long off;
      if (!CAT)
        off = ok ? (pimg + (long long)yy * W + xx) * Cin + ci : 0;
      else if (up)
        off = ok ? (pimgh + (long long)(yy >> 1) * (W >> 1) + (xx >> 1)) * C1 + ci : 0;
      else
        off = ok ? (pimg + (long long)yy * W + xx) * (Cin - C1) + (ci - C1) : 0;
      a[t] = *reinterpret_cast<const uint4*>((CAT && !up ? x2 : x) + off);
      msk[t] = ok ? 0xFFFFFFFFu : 0u;
      b[t] = *reinterpret_cast<const uint4*>(
          WT ? w + ((long long)(tap * cps + sblk) * Cout + n_base + r) * 16 + kb * 8
             : w + ((long long)(n_base + r) * 9 + tap) * Cin + ci);
    }
#pragma unroll
    for (int t = 0; t < KS_MAX; ++t) {
      if (t >= np) continue;
      const bf16x8 af = __builtin_bit_cast(
          bf16x8,
          make_uint4(a[t].x & msk[t], a[t].y & msk[t], a[t].z & msk[t], a[t].w & msk[t]));
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, __builtin_bit_cast(bf16x8, b[t]), acc,
                                                    0, 0, 0);
    }
  }
  // lane l holds column l mod 32 and rows 8 (e / 4) + 4 (l / 32) + e mod 4
  float* mine = ks_part + z * 1024;
#pragma unroll
  for (int e = 0; e < 16; ++e) mine[(8 * (e >> 2) + 4 * kb + (e & 3)) * 32 + r] = acc[e];
  __syncthreads();
  float* outv = ks_part + S * 1024;  // POOL: the tile's bf16 outputs
  for (int o = threadIdx.x; o < 1024; o += 64 * S) {
    const int row = o >> 5, col = o & 31;
    float v = 0.0f;
    for (int q = 0; q < S; ++q) v += ks_part[q * 1024 + o];
    v += bias ? bias[n_base + col] : 0.0f;
    if (relu) v = fmaxf(v, 0.0f);
    const long long mo = pixel_of(row);
    const uint16_t ob = tobf(v);
    if (mo < M) y[mo * Cout + n_base + col] = ob;
    if (POOL) outv[o] = bf(ob);
  }
  if (POOL) {
    __syncthreads();
    // pooled pixel (ty, 8 tx + q) from tile rows 2q, 2q + 1 (top) and
    // 16 + 2q, 17 + 2q (bottom)
    for (int o = threadIdx.x; o < 256; o += 64 * S) {
      const int q = o >> 5, col = o & 31;
      const float sum = ((outv[(2 * q) * 32 + col] + outv[(2 * q + 1) * 32 + col]) +
                         outv[(16 + 2 * q) * 32 + col]) +
                        outv[(17 + 2 * q) * 32 + col];
      const long long po = (ti * (H >> 1) + ty) * (long long)(W >> 1) + 8 * tx + q;
      yp[po * Cout + n_base + col] = tobf(sum * 0.25f);
    }
  }
}

// w[co][tap][ci] -> w'[ci][8 - tap][co]
__global__ void k_flip_weights(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                               int Cin, int Cout) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= Cout * 9 * Cin) return;
  const int ci = i % Cin;
  const int tap = (i / Cin) % 9;
  const int co = i / (9 * Cin);
  wt[((long long)ci * 9 + (8 - tap)) * Cout + co] = w[i];
}

// ---------------------------------------------------------------- wgrad
// dw[co][tap][ci] = sum_m dy[m][co] * x[m shifted by tap][ci]
// grid: (Cout/64 * Cin/64, 9, splits); block 256 (2x2 waves of 32x32)
constexpr int WG_CHUNK = 64;

__global__ void __launch_bounds__(256)
    k_conv3x3_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                    float* __restrict__ part, int N, int H, int W, int Cin,
                    int Cout, long long pix_per_split) {
  __shared__ uint16_t sA[64][WG_CHUNK + 8];  // [co][pixel]
  __shared__ uint16_t sB[64][WG_CHUNK + 8];  // [ci][pixel]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, kb = lane >> 5;
  const int tiles_ci = (Cin + 63) / 64;
  const int co0 = (blockIdx.x / tiles_ci) * 64;
  const int ci0 = (blockIdx.x % tiles_ci) * 64;
  const int tap = blockIdx.y;
  const int dy_ = tap / 3 - 1, dx_ = tap % 3 - 1;
  const long long M = (long long)N * H * W;
  const long long p_begin = (long long)blockIdx.z * pix_per_split;
  const long long p_end = p_begin + pix_per_split < M ? p_begin + pix_per_split : M;
  const int wm = wave >> 1, wn = wave & 1;  // wave's 32x32 quadrant
  f32x16 acc;
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
  const int hw = H * W;
  for (long long p0 = p_begin; p0 < p_end; p0 += WG_CHUNK) {
    // stage: 64 pixels x 64 channels of dy and of the shifted x, 8
    // channels (one 16-byte load) per thread and pass, zero beyond C
    for (int it = 0; it < 2; ++it) {
      const int idx = threadIdx.x + it * 256;
      const int pl = idx >> 3;         // pixel within chunk (0..63)
      const int cq = (idx & 7) * 8;    // channel offset (0..56)
      const long long m = p0 + pl;
      uint4 va = make_uint4(0, 0, 0, 0), vb = va;
      if (m < p_end) {
        if (co0 + cq < Cout)
          va = *reinterpret_cast<const uint4*>(dy + m * Cout + co0 + cq);
        const long long img = m / hw;
        const int rem = (int)(m - img * hw);
        const int yy = rem / W + dy_, xx = rem % W + dx_;
        if (ci0 + cq < Cin && yy >= 0 && yy < H && xx >= 0 && xx < W)
          vb = *reinterpret_cast<const uint4*>(
              x + (img * hw + (long long)yy * W + xx) * Cin + ci0 + cq);
      }
      const uint16_t* ta = reinterpret_cast<const uint16_t*>(&va);
      const uint16_t* tb = reinterpret_cast<const uint16_t*>(&vb);
      for (int t = 0; t < 8; ++t) {
        sA[cq + t][pl] = ta[t];
        sB[cq + t][pl] = tb[t];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k0 = 0; k0 < WG_CHUNK; k0 += 16) {
      const bf16x8 a = ld_frag(&sA[wm * 32 + r][k0 + kb * 8]);
      const bf16x8 b = ld_frag(&sB[wn * 32 + r][k0 + kb * 8]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // partial[split][co][tap][ci]
  const long long base = (long long)blockIdx.z * Cout * 9 * Cin;
  for (int e = 0; e < 16; ++e) {
    const int co = co0 + wm * 32 + 8 * (e >> 2) + 4 * kb + (e & 3);
    const int ci = ci0 + wn * 32 + r;
    if (co < Cout && ci < Cin) part[base + ((long long)co * 9 + tap) * Cin + ci] = acc[e];
  }
}

// out[i] = sum_s part[s*n + i], fixed order
__global__ void k_sum_splits(const float* __restrict__ part, float* __restrict__ out,
                             long long n, int splits) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += part[(long long)k * n + i];
  out[i] = s;
}

// ---------------------------------------------------------------- elementwise
__global__ void k_relu_bwd(const uint16_t* __restrict__ y, const uint16_t* __restrict__ dy,
                           uint16_t* __restrict__ dz, long long n) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  dz[i] = bf(y[i]) > 0.0f ? dy[i] : (uint16_t)0;
}

// per-channel column sums of a (rows x C) bf16 matrix: partials per block
__global__ void k_colsum(const uint16_t* __restrict__ a, float* __restrict__ part,
                         long long rows, int C, long long rows_per_block) {
  const long long r0 = blockIdx.x * rows_per_block;
  const long long r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  for (int c = threadIdx.x; c < C; c += 256) {
    float s = 0.0f;
    for (long long r = r0; r < r1; ++r) s += bf(a[r * C + c]);
    part[(long long)blockIdx.x * C + c] = s;
  }
}

__global__ void k_avgpool2(const uint16_t* __restrict__ x, uint16_t* __restrict__ y,
                           int N, int H, int W, int C) {  // H, W: input size
  const int Ho = H / 2, Wo = W / 2;
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= (long long)N * Ho * Wo * C) return;
  const int c = i % C;
  const long long p = i / C;
  const int xo = p % Wo, yo = (p / Wo) % Ho;
  const long long n = p / ((long long)Wo * Ho);
  const long long b = ((n * H + 2 * yo) * W + 2 * xo) * C + c;
  const float s = ((bf(x[b]) + bf(x[b + C])) + bf(x[b + (long long)W * C])) +
                  bf(x[b + (long long)W * C + C]);
  y[i] = tobf(s * 0.25f);
}

__global__ void k_avgpool2_bwd(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx,
                               int N, int H, int W, int C) {  // H, W: input size
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= (long long)N * H * W * C) return;
  const int c = i % C;
  const long long p = i / C;
  const int xx = p % W, yy = (p / W) % H;
  const long long n = p / ((long long)W * H);
  dx[i] = tobf(bf(dy[((n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c]) * 0.25f);
}

__global__ void k_upsample2(const uint16_t* __restrict__ x, uint16_t* __restrict__ y,
                            int N, int H, int W, int C) {  // H, W: output size
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= (long long)N * H * W * C) return;
  const int c = i % C;
  const long long p = i / C;
  const int xx = p % W, yy = (p / W) % H;
  const long long n = p / ((long long)W * H);
  y[i] = x[((n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
}

__global__ void k_upsample2_bwd(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx,
                                int N, int H, int W, int C) {  // H, W: output size
  const int Hi = H / 2, Wi = W / 2;
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= (long long)N * Hi * Wi * C) return;
  const int c = i % C;
  const long long p = i / C;
  const int xi = p % Wi, yi = (p / Wi) % Hi;
  const long long n = p / ((long long)Wi * Hi);
  const long long b = ((n * H + 2 * yi) * W + 2 * xi) * C + c;
  const float s = ((bf(dy[b]) + bf(dy[b + C])) + bf(dy[b + (long long)W * C])) +
                  bf(dy[b + (long long)W * C + C]);
  dx[i] = tobf(s);
}

// y[p][0:Ca] = a[p], y[p][Ca:Ca+Cb] = b[p]; backward with split = 1
__global__ void k_concat(uint16_t* __restrict__ y, uint16_t* __restrict__ a,
                         uint16_t* __restrict__ b, long long P, int Ca, int Cb,
                         int split) {
  const int C = Ca + Cb;
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= P * C) return;
  const int c = i % C;
  const long long p = i / C;
  if (!split) {
    y[i] = c < Ca ? a[p * Ca + c] : b[p * Cb + (c - Ca)];
  } else {
    if (c < Ca) a[p * Ca + c] = y[i]; else b[p * Cb + (c - Ca)] = y[i];
  }
}

// 1x1 head: out[p] = sum_c x[p][c] w[c] + bias (float32 out)
// scal != nullptr: the output is then multiplied by scal[0] (the learned
// correction's rescaling, one rounding after the head's, as a separate pass
// would do it)
__global__ void k_head(const uint16_t* __restrict__ x, const float* __restrict__ w,
                       const float* __restrict__ bias, float* __restrict__ out,
                       long long P, int C, const float* __restrict__ scal = nullptr) {
  const long long p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= P) return;
  float s = 0.0f;
  for (int c = 0; c < C; ++c) s += bf(x[p * C + c]) * w[c];
  const float v = s + bias[0];
  out[p] = scal ? v * scal[0] : v;
}

// head backward: dx[p][c] = dout[p] w[c]; per-block partials of
// dw[c] = sum_p dout[p] x[p][c] and db = sum_p dout[p] in part[blk][C+1]
__global__ void k_head_bwd(const uint16_t* __restrict__ x, const float* __restrict__ w,
                           const float* __restrict__ dout, uint16_t* __restrict__ dx,
                           float* __restrict__ part, long long P, int C,
                           long long rows_per_block) {
  const long long r0 = blockIdx.x * rows_per_block;
  const long long r1 = r0 + rows_per_block < P ? r0 + rows_per_block : P;
  for (long long i = r0 * C + threadIdx.x; i < r1 * C; i += 256) {
    const long long p = i / C;
    dx[i] = tobf(dout[p] * w[i % C]);
  }
  for (int c = threadIdx.x; c <= C; c += 256) {
    float s = 0.0f;
    for (long long p = r0; p < r1; ++p) s += c < C ? dout[p] * bf(x[p * C + c]) : dout[p];
    part[(long long)blockIdx.x * (C + 1) + c] = s;
  }
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ------------------------------------------------------------------ C-ABI
#define NX_U(stream) ((hipStream_t)(stream))
#define NX_UNET_RET()                                         \
  do {                                                        \
    hipError_t e_ = hipGetLastError();                        \
    return e_ == hipSuccess ? NX_OK                           \
                            : nx::set_error(NX_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

namespace nx {
int set_error(int code, const std::string& msg);
}

extern "C" {

// ypool != nullptr: also the 2 x 2 average pool of y (k_avgpool2's
// values), fused into the in-workgroup split-K convolution's epilogue when
// that runs and the size allows, else by k_avgpool2 after the convolution.
static int conv_fwd(const void* x, const void* w, const float* bias, void* y, int n,
                    int h, int wd, int cin, int cout, int relu, float* part, int splits,
                    hipStream_t s, bool tiled = false, const void* x2 = nullptr,
                    int c1 = 0, void* ypool = nullptr) {
  const long long M = (long long)n * h * wd;
  const int steps = 9 * cin / 16;
  const int per = part ? (steps + splits - 1) / splits : 0;
  const int used = part ? (steps + per - 1) / per : 1;
  const uint16_t* xs = (const uint16_t*)x;
  const uint16_t* ws = (const uint16_t*)w;
  uint16_t* ys = (uint16_t*)y;
  const uint16_t* x2s = (const uint16_t*)x2;
  if (part && used <= KS_WAVES && per <= KS_MAX * KS_PHASES) {
    // every split fits one wave of a workgroup: in-workgroup split-K, one
    // load phase when the splits take at most KS_MAX steps (fewer
    // registers), else two
    const dim3 grid((unsigned)((M + 31) / 32), cout / 32);
    const bool pool = ypool && wd % 16 == 0 && h % 2 == 0 && used < KS_WAVES;
    const size_t lds = (size_t)(used + (pool ? 1 : 0)) * 1024 * sizeof(float);
    uint16_t* yp = (uint16_t*)ypool;
#define NX_KS_LAUNCH(PH, PL)                                                                \
  if (x2s)                                                                                  \
    k_conv3x3_ks<true, true, PH, PL><<<grid, 64 * used, lds, s>>>(                          \
        xs, ws, bias, ys, n, h, wd, cin, cout, relu, per, steps, x2s, c1, yp);              \
  else if (tiled)                                                                           \
    k_conv3x3_ks<true, false, PH, PL><<<grid, 64 * used, lds, s>>>(                         \
        xs, ws, bias, ys, n, h, wd, cin, cout, relu, per, steps, nullptr, 0, yp);           \
  else                                                                                      \
    k_conv3x3_ks<false, false, PH, PL><<<grid, 64 * used, lds, s>>>(                        \
        xs, ws, bias, ys, n, h, wd, cin, cout, relu, per, steps, nullptr, 0, yp);
    if (per <= KS_MAX) {
      if (pool) {
        NX_KS_LAUNCH(1, true)
      } else {
        NX_KS_LAUNCH(1, false)
      }
    } else {
      NX_KS_LAUNCH(KS_PHASES, false)
    }
#undef NX_KS_LAUNCH
    if (ypool && !(pool && per <= KS_MAX))
      k_avgpool2<<<blocks_for((long long)n * (h / 2) * (wd / 2) * cout), 256, 0, s>>>(
          ys, yp, n, h, wd, cout);
    return 0;
  }
  if (x2s) {
    // decoder convolution on the upsampled + concatenated input (tiled weights)
    if (cout % 64 == 0) {
      dim3 grid((unsigned)((M + 4 * 64 - 1) / (4 * 64)), cout / 64, used);
      k_conv3x3<2, 2, true, true><<<grid, 256, 0, s>>>(xs, ws, bias, ys, n, h, wd, cin, cout,
                                                       relu, part, per, x2s, c1);
    } else {
      dim3 grid((unsigned)((M + 4 * 64 - 1) / (4 * 64)), cout / 32, used);
      k_conv3x3<2, 1, true, true><<<grid, 256, 0, s>>>(xs, ws, bias, ys, n, h, wd, cin, cout,
                                                       relu, part, per, x2s, c1);
    }
  } else if (cout % 64 == 0) {
    dim3 grid((unsigned)((M + 4 * 64 - 1) / (4 * 64)), cout / 64, used);
    if (tiled)
      k_conv3x3<2, 2, true><<<grid, 256, 0, s>>>(xs, ws, bias, ys, n, h, wd, cin, cout, relu,
                                                 part, per);
    else
      k_conv3x3<2, 2, false><<<grid, 256, 0, s>>>(xs, ws, bias, ys, n, h, wd, cin, cout, relu,
                                                  part, per);
  } else {
    dim3 grid((unsigned)((M + 4 * 64 - 1) / (4 * 64)), cout / 32, used);
    if (tiled)
      k_conv3x3<2, 1, true><<<grid, 256, 0, s>>>(xs, ws, bias, ys, n, h, wd, cin, cout, relu,
                                                 part, per);
    else
      k_conv3x3<2, 1, false><<<grid, 256, 0, s>>>(xs, ws, bias, ys, n, h, wd, cin, cout, relu,
                                                  part, per);
  }
  if (part) {
    const long long mc = M * cout;
    k_conv_finish<<<(unsigned)((mc + 255) / 256), 256, 0, s>>>(part, bias, (uint16_t*)y, mc,
                                                               cout, used, relu);
  }
  if (ypool)
    k_avgpool2<<<blocks_for((long long)n * (h / 2) * (wd / 2) * cout), 256, 0, s>>>(
        ys, (uint16_t*)ypool, n, h, wd, cout);
  return 0;
}

// w[co][tap][ci] (KRSC) -> wt[tap][ci / 16][co][16] for the WT convolution
__global__ void k_tile_weights(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                               int cin, int cout) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= 9LL * cin * cout) return;
  const int e = (int)(i % 16);
  long long t = i / 16;
  const int co = (int)(t % cout);
  t /= cout;
  const int s = (int)(t % (cin / 16)), tap = (int)(t / (cin / 16));
  wt[i] = w[((long long)co * 9 + tap) * cin + 16 * s + e];
}

int nx_conv3x3_fwd(const void* x, const void* w, const float* bias, void* y, int n,
                   int h, int wd, int cin, int cout, int relu, void* stream) {
  if (!x || !w || !y || n < 1 || h < 1 || wd < 1 || cin < 8 || cin % 8 || cout % 32)
    return nx::set_error(NX_ERR_ARG, "conv3x3: channels must be cin % 8 == 0, cout % 32 == 0");
  conv_fwd(x, w, bias, y, n, h, wd, cin, cout, relu, nullptr, 1, NX_U(stream));
  NX_UNET_RET();
}

int nx_conv3x3_fwd_split(const void* x, const void* w, const float* bias, void* y,
                         float* work, int splits, int n, int h, int wd, int cin, int cout,
                         int relu, void* stream) {
  if (!x || !w || !y || !work || n < 1 || h < 1 || wd < 1 || cin % 16 || cout % 32 ||
      splits < 1 || splits > 9 * cin / 16)
    return nx::set_error(NX_ERR_ARG,
                         "conv3x3 split: cin % 16 == 0, cout % 32 == 0, 1 <= splits <= 9 cin / 16");
  conv_fwd(x, w, bias, y, n, h, wd, cin, cout, relu, work, splits, NX_U(stream));
  NX_UNET_RET();
}

int nx_conv3x3_dgrad(const void* dy, const void* w, void* dx, void* wt, int n, int h,
                     int wd, int cin, int cout, void* stream) {
  if (!dy || !w || !dx || !wt || cin % 32 || cout % 16)
    return nx::set_error(NX_ERR_ARG, "conv3x3 dgrad: cin % 32 == 0 and cout % 16 == 0");
  const hipStream_t s = NX_U(stream);
  k_flip_weights<<<blocks_for((long long)cout * 9 * cin), 256, 0, s>>>(
      (const uint16_t*)w, (uint16_t*)wt, cin, cout);
  if (hipGetLastError() != hipSuccess) return nx::set_error(NX_ERR_HIP, "flip failed");
  return nx_conv3x3_fwd(dy, wt, nullptr, dx, n, h, wd, cout, cin, 0, stream);
}

// dw: float32 [cout][9][cin]; work: float32, splits * cout * 9 * cin
int nx_conv3x3_wgrad(const void* x, const void* dy, float* dw, float* work, int splits,
                     int n, int h, int wd, int cin, int cout, void* stream) {
  if (!x || !dy || !dw || !work || splits < 1 || cin % 8 || cout % 8)
    return nx::set_error(NX_ERR_ARG, "conv3x3 wgrad: cin % 8 == 0 and cout % 8 == 0");
  const long long M = (long long)n * h * wd;
  long long per = (M + splits - 1) / splits;
  per = (per + WG_CHUNK - 1) / WG_CHUNK * WG_CHUNK;
  const int used = (int)((M + per - 1) / per);
  const hipStream_t s = NX_U(stream);
  dim3 grid(((cout + 63) / 64) * ((cin + 63) / 64), 9, used);
  k_conv3x3_wgrad<<<grid, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)dy, work, n,
                                       h, wd, cin, cout, per);
  if (hipGetLastError() != hipSuccess) return nx::set_error(NX_ERR_HIP, "wgrad failed");
  const long long nw = (long long)cout * 9 * cin;
  k_sum_splits<<<blocks_for(nw), 256, 0, s>>>(work, dw, nw, used);
  NX_UNET_RET();
}

int nx_relu_bwd(const void* y, const void* dy, void* dz, long long count, void* stream) {
  k_relu_bwd<<<blocks_for(count), 256, 0, NX_U(stream)>>>(
      (const uint16_t*)y, (const uint16_t*)dy, (uint16_t*)dz, count);
  NX_UNET_RET();
}

// column sums of a bf16 (rows x c) matrix into out[c] (float32);
// work: float32 of blocks * c, blocks = min(1024, rows)
int nx_colsum(const void* a, float* out, float* work, long long rows, int c,
              void* stream) {
  const long long blocks = rows < 1024 ? rows : 1024;
  const long long per = (rows + blocks - 1) / blocks;
  const hipStream_t s = NX_U(stream);
  k_colsum<<<(unsigned)blocks, 256, 0, s>>>((const uint16_t*)a, work, rows, c, per);
  if (hipGetLastError() != hipSuccess) return nx::set_error(NX_ERR_HIP, "colsum failed");
  k_sum_splits<<<blocks_for(c), 256, 0, s>>>(work, out, c, (int)blocks);
  NX_UNET_RET();
}

int nx_avgpool2(const void* x, void* y, int n, int h, int wd, int c, int backward,
                void* stream) {
  const long long cnt = (long long)n * (backward ? h * wd : (h / 2) * (wd / 2)) * c;
  if (backward)
    k_avgpool2_bwd<<<blocks_for(cnt), 256, 0, NX_U(stream)>>>(
        (const uint16_t*)x, (uint16_t*)y, n, h, wd, c);
  else
    k_avgpool2<<<blocks_for(cnt), 256, 0, NX_U(stream)>>>(
        (const uint16_t*)x, (uint16_t*)y, n, h, wd, c);
  NX_UNET_RET();
}

int nx_upsample2(const void* x, void* y, int n, int h, int wd, int c, int backward,
                 void* stream) {
  const long long cnt = (long long)n * (backward ? (h / 2) * (wd / 2) : h * wd) * c;
  if (backward)
    k_upsample2_bwd<<<blocks_for(cnt), 256, 0, NX_U(stream)>>>(
        (const uint16_t*)x, (uint16_t*)y, n, h, wd, c);
  else
    k_upsample2<<<blocks_for(cnt), 256, 0, NX_U(stream)>>>(
        (const uint16_t*)x, (uint16_t*)y, n, h, wd, c);
  NX_UNET_RET();
}

int nx_concat(void* y, void* a, void* b, long long pixels, int ca, int cb, int split,
              void* stream) {
  k_concat<<<blocks_for(pixels * (ca + cb)), 256, 0, NX_U(stream)>>>(
      (uint16_t*)y, (uint16_t*)a, (uint16_t*)b, pixels, ca, cb, split);
  NX_UNET_RET();
}

int nx_head_fwd(const void* x, const float* w, const float* bias, float* out,
                long long pixels, int c, void* stream) {
  k_head<<<blocks_for(pixels), 256, 0, NX_U(stream)>>>((const uint16_t*)x, w, bias, out,
                                                      pixels, c);
  NX_UNET_RET();
}

// dw[c] (c floats) and db (1 float) into dwb[0..c]; work: 1024 * (c+1) floats
int nx_head_bwd(const void* x, const float* w, const float* dout, void* dx, float* dwb,
                float* work, long long pixels, int c, void* stream) {
  const long long blocks = pixels < 1024 ? pixels : 1024;
  const long long per = (pixels + blocks - 1) / blocks;
  const hipStream_t s = NX_U(stream);
  k_head_bwd<<<(unsigned)blocks, 256, 0, s>>>((const uint16_t*)x, w, dout, (uint16_t*)dx,
                                              work, pixels, c, per);
  if (hipGetLastError() != hipSuccess) return nx::set_error(NX_ERR_HIP, "head bwd failed");
  k_sum_splits<<<blocks_for(c + 1), 256, 0, s>>>(work, dwb, c + 1, (int)blocks);
  NX_UNET_RET();
}

}  // extern "C"

namespace nx {
// the native U-Net runner's convolutions on tiled weights (k_tile_weights);
// same arithmetic, results and arguments as nx_conv3x3_fwd(_split) at n = 1
int conv3x3_tile_weights(const void* w, void* wt, int cin, int cout, hipStream_t s) {
  if (!w || !wt || cin % 16 || cout % 32)
    return set_error(NX_ERR_ARG, "weight tiling: cin % 16 == 0, cout % 32 == 0");
  const long long n = 9LL * cin * cout;
  k_tile_weights<<<(unsigned)((n + 255) / 256), 256, 0, s>>>((const uint16_t*)w,
                                                             (uint16_t*)wt, cin, cout);
  NX_UNET_RET();
}

// nx_head_fwd with the output multiplied by scal[0] (device scalar)
int head_fwd_scaled(const void* x, const float* w, const float* bias, float* out,
                    long long pixels, int c, const float* scal, hipStream_t s) {
  if (!x || !w || !bias || !out || !scal || pixels < 1 || c < 1)
    return set_error(NX_ERR_ARG, "head: null argument or empty shape");
  k_head<<<blocks_for(pixels), 256, 0, s>>>((const uint16_t*)x, w, bias, out, pixels, c, scal);
  NX_UNET_RET();
}

int conv3x3_fwd_tiled(const void* x, const void* wt, const float* bias, void* y, float* work,
                      int splits, int h, int wd, int cin, int cout, int relu, hipStream_t s,
                      const void* x2, int c1, void* ypool) {
  if (!x || !wt || !y || h < 1 || wd < 1 || cin % 16 || cout % 32 || splits < 1 ||
      splits > 9 * cin / 16 || (splits > 1 && !work))
    return set_error(NX_ERR_ARG, "tiled conv3x3: cin % 16 == 0, cout % 32 == 0, valid splits");
  if (x2 && (c1 < 16 || c1 % 16 || c1 >= cin || h % 2 || wd % 2))
    return set_error(NX_ERR_ARG, "concatenating conv3x3: 16 <= c1 < cin, c1 % 16 == 0, even size");
  if (ypool && (h % 2 || wd % 2))
    return set_error(NX_ERR_ARG, "pooled conv3x3: even size");
  conv_fwd(x, wt, bias, y, 1, h, wd, cin, cout, relu, splits > 1 ? work : nullptr, splits, s,
           true, x2, c1, ypool);
  NX_UNET_RET();
}
}  // namespace nx
