// Native runner of the learned coarse correction (inference only).
//
// One call of nx_learned_correct enqueues, on the simulator's stream and
// with no host synchronisation, the whole correction of one learned cycle:
//   1. restriction of the current residual to the coarse grid
//      (with Chebyshev coarse smoothing: rc = the block sum of the
//      residual and the coarse faces' open fractions, the pair of the
//      face-weighted coarse operator; with Jacobi: rc = f^2 x fluid mean),
//      coarse mask; the network's grid may extend
//      beyond ceil(W / f) x ceil(H / f), the extra cells solid);
//   2. k_pack_input (one block): s = RMS of rc over fluid coarse cells,
//      network input x[p] = (bf16(rc/s), bf16(solid), 0 x 6);
//   3. the U-Net forward (4 levels, 3x3 conv + ReLU pairs, 2x2 average pool,
//      2x nearest upsample + channel concat, 1x1 head) with the forward
//      convolution kernels, split over K exactly as the Python wrapper
//      splits them, so nx_unet_forward equals the Python forward bitwise
//      (the decoder convolutions read the upsampled, concatenated input in
//      their fragment loads: the same values as the Python path's copies);
//   4. e = s x head output (applied by the head kernel);
//   5. Chebyshev smoothing of the coarse problem from e (the face-weighted
//      operator, nx_coarse_smooth_faces);
//   6. k_center_output (one block): e -= fluid mean of e, 0 on solid cells;
//   7. interpolation and the energy line search (nx_prolong_line_search).
// Steps 2-6 touch only the runner's buffers; from the second call with the
// same stream and smoothing parameters they replay as one HIP graph (no
// memset nodes: every buffer in them is fully written by a kernel).
// The weights stay owned by the caller (device pointers handed to
// nx_unet_create); the runner owns its activation buffers, sized for one
// coarse grid.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/natrix.h"
#include "nx_common.h"
#include "nx_kernels.h"

struct nx_unet {
  int H = 0, W = 0, base = 0;
  // 14 convolutions in forward order: e0 a/b, e1 a/b, e2 a/b, e3 a/b,
  // d2 a/b, d1 a/b, d0 a/b
  const void* w[14] = {};
  void* wt[14] = {};  // tiled copies (cin % 16 == 0), owned by the runner
  const float* b[14] = {};
  int cin[14] = {}, cout[14] = {};
  const float* head_w = nullptr;
  const float* head_b = nullptr;
  std::vector<void*> bufs;
  uint16_t *x0 = nullptr, *t0 = nullptr, *e0 = nullptr, *p0 = nullptr;
  uint16_t *t1 = nullptr, *e1 = nullptr, *p1 = nullptr;
  uint16_t *t2 = nullptr, *e2 = nullptr, *p2 = nullptr;
  uint16_t *t3 = nullptr, *e3 = nullptr;
  // decoder: s = first conv (on the upsampled + concatenated input, read in
  // place), d = second conv
  uint16_t *s2 = nullptr, *d2 = nullptr;
  uint16_t *s1 = nullptr, *d1 = nullptr;
  uint16_t *s0 = nullptr, *d0 = nullptr;
  float* out = nullptr;    // head output / coarse correction e
  float* rc = nullptr;     // coarse right-hand side
  uint8_t* mc = nullptr;   // coarse mask
  float* work = nullptr;   // split-K partial sums
  size_t work_elems = 0;
  float* scal = nullptr;   // [0] = s (RMS of rc over fluid cells)
  float* cws = nullptr;    // Chebyshev workspace, 6 H W floats
  float* faces = nullptr;  // coarse face weights, 2 H W floats (east, south)
  // three-level cycle (mid_f > 1): a middle grid mid_f times coarser than
  // the fine one, smoothed with mid_iters Chebyshev iterations on
  // [mid_lmin, 8] of its face-weighted operator; buffers for Wm x Hm
  int mid_f = 0, mid_iters = 0;
  float mid_lmin = 0.0f;
  int Wm = 0, Hm = 0;
  float *rm = nullptr, *em = nullptr, *mfaces = nullptr, *mws = nullptr;
  uint8_t* mm = nullptr;
  // steps 2-6 captured as one HIP graph (every pointer in them is the
  // runner's own), keyed by the stream and the smoothing parameters; the
  // capture happens on the second call with the same key, when everything
  // the pipeline touches has been allocated and run once eagerly
  hipGraphExec_t graph = nullptr;
  hipStream_t g_stream = nullptr, w_stream = nullptr;
  int g_iters = -1, w_iters = -1;
  float g_lmin = 0.0f, w_lmin = 0.0f;
  // bypass != 0: the coarse correction starts from zero instead of the
  // network's output (the same cycle without the network, for comparison)
  int bypass = 0;
  int g_bypass = -1, w_bypass = -1;
};

namespace {

int fail(int code, const char* msg) { return nx::set_error(code, msg); }

// Same rule as natrix_amd.learned.unet.conv_splits.  A split count whose
// splits do not fit one workgroup's waves (at most 16 splits of at most 9
// K steps, one load phase of the in-workgroup split-K convolution) is
// replaced by the smallest one that does, else by 16 splits when those
// take at most 18 steps (two load phases).
int conv_splits(long long m, int cin, int cout) {
  if (cin % 16) return 1;
  const long long base = ((m + 255) / 256) * (cout / (cout % 64 == 0 ? 64 : 32));
  if (base >= 256) return 1;
  const int steps = 9 * cin / 16;
  const long long a = steps / 8, b = (256 + base - 1) / base;
  long long s = a < b ? a : b;
  if (s <= 1) return 1;
  const long long per = (steps + s - 1) / s, used = (steps + per - 1) / per;
  if (used > 16 || per > 9) {
    const long long s2 = (steps + 8) / 9;
    if (s2 <= 16)
      s = s2;
    else if ((steps + 15) / 16 <= 18)
      s = 16;
  }
  return (int)s;
}

__device__ __forceinline__ float block_sum_1024(float v, float* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  float t = 0.0f;
  for (int w = 0; w < 16; ++w) t += sh[w];
  __syncthreads();
  return t;
}

// x[p] = (bf16(rc[p] / s), bf16(mask), 0, ...), s = sqrt(sum_fluid rc^2 / n)
// The coarse grid has at most 128 x 128 cells: each of the 1024 threads
// holds its (at most) 16 cells i = t, t + 1024, ... in registers, so the
// loads of a pass are issued together; sums run over them in that order.
constexpr int CELLS_PER_THREAD = 16;

// x[p] = (bf16(rc[p] / s), bf16(mask), 0, ...), s = sqrt(sum_fluid rc^2 / n)
__global__ void __launch_bounds__(1024)
    k_pack_input(const float* __restrict__ rc, const uint8_t* __restrict__ mc,
                 uint16_t* __restrict__ x, float* __restrict__ scal, int n) {
  __shared__ float sh[16];
  float v[CELLS_PER_THREAD];
  bool solid[CELLS_PER_THREAD];
#pragma unroll
  for (int k = 0; k < CELLS_PER_THREAD; ++k) {
    const int i = threadIdx.x + 1024 * k;
    v[k] = i < n ? rc[i] : 0.0f;
    solid[k] = i < n ? mc[i] != 0 : true;
  }
  float ss = 0.0f, cnt = 0.0f;
#pragma unroll
  for (int k = 0; k < CELLS_PER_THREAD; ++k)
    if (!solid[k]) {
      ss += v[k] * v[k];
      cnt += 1.0f;
    }
  ss = block_sum_1024(ss, sh);
  cnt = block_sum_1024(cnt, sh);
  const float s = fmaxf(sqrtf(ss / fmaxf(cnt, 1.0f)), 1e-30f);
  if (threadIdx.x == 0) scal[0] = s;
#pragma unroll
  for (int k = 0; k < CELLS_PER_THREAD; ++k) {
    const int i = threadIdx.x + 1024 * k;
    if (i >= n) break;
    const uint32_t lo = nx::f32_to_bf16(v[k] / s) | (nx::f32_to_bf16(solid[k] ? 1.0f : 0.0f) << 16);
    reinterpret_cast<uint4*>(x)[i] = make_uint4(lo, 0u, 0u, 0u);
  }
}

__global__ void k_zero_output(float* __restrict__ e, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) e[i] = 0.0f;
}

// e -= mean of e over fluid cells; e = 0 on solid cells
__global__ void __launch_bounds__(1024)
    k_center_output(float* __restrict__ e, const uint8_t* __restrict__ mc, int n) {
  __shared__ float sh[16];
  float v[CELLS_PER_THREAD];
  bool solid[CELLS_PER_THREAD];
#pragma unroll
  for (int k = 0; k < CELLS_PER_THREAD; ++k) {
    const int i = threadIdx.x + 1024 * k;
    v[k] = i < n ? e[i] : 0.0f;
    solid[k] = i < n ? mc[i] != 0 : true;
  }
  float sum = 0.0f, cnt = 0.0f;
#pragma unroll
  for (int k = 0; k < CELLS_PER_THREAD; ++k)
    if (!solid[k]) {
      sum += v[k];
      cnt += 1.0f;
    }
  sum = block_sum_1024(sum, sh);
  cnt = block_sum_1024(cnt, sh);
  const float m = sum / fmaxf(cnt, 1.0f);
#pragma unroll
  for (int k = 0; k < CELLS_PER_THREAD; ++k) {
    const int i = threadIdx.x + 1024 * k;
    if (i < n) e[i] = solid[k] ? 0.0f : v[k] - m;
  }
}

// ypool: also the 2 x 2 average pool of the output (fused into the
// convolution where it can be)
int conv(nx_unet* u, int i, const void* x, void* y, int h, int w, hipStream_t s,
         void* ypool = nullptr) {
  const long long m = (long long)h * w;
  const int sp = conv_splits(m, u->cin[i], u->cout[i]);
  if (u->wt[i])
    return nx::conv3x3_fwd_tiled(x, u->wt[i], u->b[i], y, u->work, sp, h, w, u->cin[i],
                                 u->cout[i], 1, s, nullptr, 0, ypool);
  if (ypool) return fail(NX_ERR_ARG, "pooled convolution needs tiled weights");
  if (sp > 1)
    return nx_conv3x3_fwd_split(x, u->w[i], u->b[i], y, u->work, sp, 1, h, w, u->cin[i],
                                u->cout[i], 1, s);
  return nx_conv3x3_fwd(x, u->w[i], u->b[i], y, 1, h, w, u->cin[i], u->cout[i], 1, s);
}

// decoder convolution i on concat(upsample2(lo) (c1 channels), skip), read
// in the convolution's fragment loads (no upsampled or concatenated copy)
int conv_cat(nx_unet* u, int i, const void* lo, const void* skip, int c1, void* y, int h, int w,
             hipStream_t s) {
  const long long m = (long long)h * w;
  const int sp = conv_splits(m, u->cin[i], u->cout[i]);
  return nx::conv3x3_fwd_tiled(lo, u->wt[i], u->b[i], y, u->work, sp, h, w, u->cin[i],
                               u->cout[i], 1, s, skip, c1);
}

// scal != nullptr: the head's output multiplied by scal[0] (the rescaling)
int forward(nx_unet* u, const uint16_t* x, float* out, hipStream_t s,
            const float* scal = nullptr) {
  const int H = u->H, W = u->W, c = u->base;
#define NX_TRY_U(expr)          \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != NX_OK) return rc_; \
  } while (0)
  NX_TRY_U(conv(u, 0, x, u->t0, H, W, s));
  NX_TRY_U(conv(u, 1, u->t0, u->e0, H, W, s, u->p0));
  NX_TRY_U(conv(u, 2, u->p0, u->t1, H / 2, W / 2, s));
  NX_TRY_U(conv(u, 3, u->t1, u->e1, H / 2, W / 2, s, u->p1));
  NX_TRY_U(conv(u, 4, u->p1, u->t2, H / 4, W / 4, s));
  NX_TRY_U(conv(u, 5, u->t2, u->e2, H / 4, W / 4, s, u->p2));
  NX_TRY_U(conv(u, 6, u->p2, u->t3, H / 8, W / 8, s));
  NX_TRY_U(conv(u, 7, u->t3, u->e3, H / 8, W / 8, s));
  NX_TRY_U(conv_cat(u, 8, u->e3, u->e2, 8 * c, u->s2, H / 4, W / 4, s));
  NX_TRY_U(conv(u, 9, u->s2, u->d2, H / 4, W / 4, s));
  NX_TRY_U(conv_cat(u, 10, u->d2, u->e1, 4 * c, u->s1, H / 2, W / 2, s));
  NX_TRY_U(conv(u, 11, u->s1, u->d1, H / 2, W / 2, s));
  NX_TRY_U(conv_cat(u, 12, u->d1, u->e0, 2 * c, u->s0, H, W, s));
  NX_TRY_U(conv(u, 13, u->s0, u->d0, H, W, s));
  if (scal)
    NX_TRY_U(nx::head_fwd_scaled(u->d0, u->head_w, u->head_b, out, (long long)H * W, c, scal,
                                 s));
  else
    NX_TRY_U(nx_head_fwd(u->d0, u->head_w, u->head_b, out, (long long)H * W, c, s));
#undef NX_TRY_U
  return NX_OK;
}

// steps 2-6 of a correction: pack, network, rescale, coarse smoothing,
// centring (only the runner's buffers)
int coarse_pipeline(nx_unet* u, int iters, float lmin, hipStream_t s) {
  const int n = u->H * u->W;
  int rc = NX_OK;
  if (u->bypass) {
    k_zero_output<<<(n + 255) / 256, 256, 0, s>>>(u->out, n);
    if (hipGetLastError() != hipSuccess) return fail(NX_ERR_HIP, "zero launch failed");
  } else {
    k_pack_input<<<1, 1024, 0, s>>>(u->rc, u->mc, u->x0, u->scal, n);
    if (hipGetLastError() != hipSuccess) return fail(NX_ERR_HIP, "pack launch failed");
    rc = forward(u, u->x0, u->out, s, u->scal);  // the head applies the rescaling
    if (rc != NX_OK) return rc;
  }
  // Chebyshev coarse smoothing takes the face-weighted operator (with the
  // block sums its restriction wrote); Jacobi the plain one
  rc = nx::launch_coarse_smooth(s, u->W, u->H, u->out, u->rc, u->mc, iters, lmin, u->cws,
                                lmin > 0.0f ? u->faces : nullptr);
  if (rc != NX_OK) return rc;
  k_center_output<<<1, 1024, 0, s>>>(u->out, u->mc, n);
  if (hipGetLastError() != hipSuccess) return fail(NX_ERR_HIP, "center launch failed");
  return NX_OK;
}

int run_pipeline(nx_unet* u, int iters, float lmin, hipStream_t s) {
  if (u->graph && u->g_stream == s && u->g_iters == iters && u->g_lmin == lmin &&
      u->g_bypass == u->bypass)
    return hipGraphLaunch(u->graph, s) == hipSuccess ? NX_OK
                                                     : fail(NX_ERR_HIP, "hipGraphLaunch failed");
  if (u->w_stream != s || u->w_iters != iters || u->w_lmin != lmin || u->w_bypass != u->bypass) {
    u->w_stream = s;
    u->w_iters = iters;
    u->w_lmin = lmin;
    u->w_bypass = u->bypass;
    return coarse_pipeline(u, iters, lmin, s);
  }
  if (u->graph) (void)hipGraphExecDestroy(u->graph);
  u->graph = nullptr;
  hipGraph_t g;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(NX_ERR_HIP, "hipStreamBeginCapture failed");
  const int rc = coarse_pipeline(u, iters, lmin, s);
  const hipError_t ec = hipStreamEndCapture(s, &g);
  if (rc != NX_OK) {
    if (ec == hipSuccess) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(NX_ERR_HIP, "hipStreamEndCapture failed");
  const hipError_t ei = hipGraphInstantiate(&u->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    u->graph = nullptr;
    return fail(NX_ERR_HIP, "hipGraphInstantiate failed");
  }
  u->g_stream = s;
  u->g_iters = iters;
  u->g_lmin = lmin;
  u->g_bypass = u->bypass;
  return hipGraphLaunch(u->graph, s) == hipSuccess ? NX_OK
                                                   : fail(NX_ERR_HIP, "hipGraphLaunch failed");
}

}  // namespace

extern "C" {

int nx_unet_create(int H, int W, int base, const void* const* conv_w,
                   const float* const* conv_b, const float* head_w, const float* head_b,
                   nx_unet** out) {
  if (!out || !conv_w || !conv_b || !head_w || !head_b)
    return fail(NX_ERR_ARG, "null argument");
  *out = nullptr;
  if (H < 8 || W < 8 || H % 8 || W % 8 || H > 128 || W > 128 || base < 16 || base % 16)
    return fail(NX_ERR_ARG,
                "U-Net grid must be a multiple of 8 in [8, 128], base a multiple of 16");
  nx_unet* u = new nx_unet();
  u->H = H;
  u->W = W;
  u->base = base;
  const int c = base;
  const int ci[14] = {8, c, c, 2 * c, 2 * c, 4 * c, 4 * c, 8 * c, 12 * c, 4 * c, 6 * c, 2 * c,
                      3 * c, c};
  const int co[14] = {c, c, 2 * c, 2 * c, 4 * c, 4 * c, 8 * c, 8 * c, 4 * c, 4 * c, 2 * c,
                      2 * c, c, c};
  const int lvl[14] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};
  for (int i = 0; i < 14; ++i) {
    if (!conv_w[i] || !conv_b[i]) {
      delete u;
      return fail(NX_ERR_ARG, "null convolution weight");
    }
    u->w[i] = conv_w[i];
    u->b[i] = conv_b[i];
    u->cin[i] = ci[i];
    u->cout[i] = co[i];
    const long long m = (long long)(H >> lvl[i]) * (W >> lvl[i]);
    const int sp = conv_splits(m, ci[i], co[i]);
    if (sp > 1 && (size_t)(sp * m * co[i]) > u->work_elems)
      u->work_elems = (size_t)sp * m * co[i];
  }
  u->head_w = head_w;
  u->head_b = head_b;
  bool ok = true;
  auto alloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, bytes < 256 ? 256 : bytes) != hipSuccess) {
      ok = false;
      return nullptr;
    }
    u->bufs.push_back(p);
    return p;
  };
  const size_t P0 = (size_t)H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64, B = 2;
  u->x0 = (uint16_t*)alloc(P0 * 8 * B);
  u->t0 = (uint16_t*)alloc(P0 * c * B);
  u->e0 = (uint16_t*)alloc(P0 * c * B);
  u->p0 = (uint16_t*)alloc(P1 * c * B);
  u->t1 = (uint16_t*)alloc(P1 * 2 * c * B);
  u->e1 = (uint16_t*)alloc(P1 * 2 * c * B);
  u->p1 = (uint16_t*)alloc(P2 * 2 * c * B);
  u->t2 = (uint16_t*)alloc(P2 * 4 * c * B);
  u->e2 = (uint16_t*)alloc(P2 * 4 * c * B);
  u->p2 = (uint16_t*)alloc(P3 * 4 * c * B);
  u->t3 = (uint16_t*)alloc(P3 * 8 * c * B);
  u->e3 = (uint16_t*)alloc(P3 * 8 * c * B);
  u->s2 = (uint16_t*)alloc(P2 * 4 * c * B);
  u->d2 = (uint16_t*)alloc(P2 * 4 * c * B);
  u->s1 = (uint16_t*)alloc(P1 * 2 * c * B);
  u->d1 = (uint16_t*)alloc(P1 * 2 * c * B);
  u->s0 = (uint16_t*)alloc(P0 * c * B);
  u->d0 = (uint16_t*)alloc(P0 * c * B);
  u->out = (float*)alloc(P0 * sizeof(float));
  u->rc = (float*)alloc(P0 * sizeof(float));
  u->mc = (uint8_t*)alloc(P0);
  u->work = (float*)alloc((u->work_elems ? u->work_elems : 1) * sizeof(float));
  u->scal = (float*)alloc(4 * sizeof(float));
  u->cws = (float*)alloc(6 * P0 * sizeof(float));
  u->faces = (float*)alloc(2 * P0 * sizeof(float));
  void* wcopy[14] = {};
  for (int i = 0; i < 14 && ok; ++i) {
    if (ci[i] % 16 == 0)
      u->wt[i] = alloc((size_t)9 * ci[i] * co[i] * B);
    else
      wcopy[i] = alloc((size_t)9 * ci[i] * co[i] * B);
  }
  if (!ok) {
    nx_unet_destroy(u);
    return fail(NX_ERR_HIP, "U-Net workspace allocation failed");
  }
  // the runner's copies of the weights, made once on the null stream:
  // tiled where the convolution takes them, as given otherwise
  for (int i = 0; i < 14; ++i) {
    int rc = NX_OK;
    if (u->wt[i]) {
      rc = nx::conv3x3_tile_weights(u->w[i], u->wt[i], ci[i], co[i], nullptr);
      u->w[i] = nullptr;  // the tiled copy replaces the caller's weights
    } else {
      if (hipMemcpyAsync(wcopy[i], u->w[i], (size_t)9 * ci[i] * co[i] * B,
                         hipMemcpyDeviceToDevice, nullptr) != hipSuccess)
        rc = fail(NX_ERR_HIP, "U-Net weight copy failed");
      u->w[i] = wcopy[i];
    }
    if (rc != NX_OK) {
      nx_unet_destroy(u);
      return rc;
    }
  }
  if (hipStreamSynchronize(nullptr) != hipSuccess) {
    nx_unet_destroy(u);
    return fail(NX_ERR_HIP, "U-Net weight tiling failed");
  }
  *out = u;
  return NX_OK;
}

int nx_unet_destroy(nx_unet* u) {
  if (!u) return NX_OK;
  if (u->graph) (void)hipGraphExecDestroy(u->graph);
  for (void* p : u->bufs) (void)hipFree(p);
  for (void* p : {(void*)u->rm, (void*)u->em, (void*)u->mfaces, (void*)u->mws, (void*)u->mm})
    if (p) (void)hipFree(p);
  delete u;
  return NX_OK;
}

int nx_unet_forward(nx_unet* u, const void* x, float* out, void* stream) {
  if (!u || !x || !out) return fail(NX_ERR_ARG, "null argument");
  return forward(u, (const uint16_t*)x, out, (hipStream_t)stream);
}

}  // extern "C"

// apply == false leaves the line-search step pending for
// nx::smooth_chebyshev_pending, which takes it as it loads the pressure
static int learned_correct_impl(nx_sim* sim, nx_unet* u, int factor, int coarse_iters,
                                float lambda_min, double* out, bool apply) {
  if (!sim || !u || factor < 1) return fail(NX_ERR_ARG, "bad learned correction arguments");
  void* sv = nullptr;
  int rc = nx_get_stream(sim, &sv);
  if (rc != NX_OK) return rc;
  const hipStream_t s = (hipStream_t)sv;
  // the coarse grid must be the network's grid
  int W = 0, H = 0;
  rc = nx_get_size(sim, &W, &H);
  if (rc != NX_OK) return rc;
  if ((W + factor - 1) / factor > u->W || (H + factor - 1) / factor > u->H)
    return fail(NX_ERR_ARG, "the network's grid is smaller than the fine grid / factor");
  if (u->mid_f > 1 && lambda_min > 0.0f) {
    // three levels: fine -> middle (block sums and faces), middle -> the
    // network's grid (block sums; faces from the fine mask at the full
    // factor), the coarse pipeline, interpolation to the middle grid, its
    // Chebyshev smoothing, interpolation to the fine grid with the line
    // search
    const int fm = u->mid_f, f2 = factor / fm;
    if (f2 < 2 || f2 * fm != factor) return fail(NX_ERR_ARG, "factor must be a multiple of the middle factor");
    const int Wm = (W + fm - 1) / fm, Hm = (H + fm - 1) / fm;
    if ((Wm + f2 - 1) / f2 > u->W || (Hm + f2 - 1) / f2 > u->H)
      return fail(NX_ERR_ARG, "the network's grid is smaller than the middle grid / factor");
    if (Wm != u->Wm || Hm != u->Hm) {
      for (void* p : {(void*)u->rm, (void*)u->em, (void*)u->mfaces, (void*)u->mws, (void*)u->mm})
        if (p) (void)hipFree(p);
      u->rm = u->em = u->mfaces = u->mws = nullptr;
      u->mm = nullptr;
      u->Wm = u->Hm = 0;
      const size_t n = (size_t)Wm * Hm;
      if (hipMalloc(&u->rm, n * 4) != hipSuccess || hipMalloc(&u->em, n * 4) != hipSuccess ||
          hipMalloc(&u->mfaces, 2 * n * 4) != hipSuccess ||
          hipMalloc(&u->mws, 6 * n * 4) != hipSuccess || hipMalloc(&u->mm, n) != hipSuccess)
        return fail(NX_ERR_HIP, "middle-level buffers");
      u->Wm = Wm;
      u->Hm = Hm;
    }
    rc = nx::restrict_residual_padded(sim, fm, Wm, Hm, u->rm, u->mm, u->mfaces);
    if (rc != NX_OK) return rc;
    rc = nx::block_sum_restrict(s, u->rm, u->mm, Wm, Hm, f2, u->rc, u->mc, u->W, u->H);
    if (rc != NX_OK) return rc;
    rc = nx::coarse_faces_padded(sim, factor, u->W, u->H, u->faces);
    if (rc != NX_OK) return rc;
    rc = run_pipeline(u, coarse_iters, lambda_min, s);
    if (rc != NX_OK) return rc;
    rc = nx::prolong_to(s, u->em, u->mm, Wm, Hm, f2, u->out, u->mc, u->W, u->H);
    if (rc != NX_OK) return rc;
    rc = nx::launch_coarse_smooth(s, Wm, Hm, u->em, u->rm, u->mm, u->mid_iters, u->mid_lmin,
                                  u->mws, u->mfaces);
    if (rc != NX_OK) return rc;
    return nx::prolong_line_search_padded(sim, fm, Wm, Hm, u->em, u->mm, out, apply);
  }
  rc = nx::restrict_residual_padded(sim, factor, u->W, u->H, u->rc, u->mc,
                                    lambda_min > 0.0f ? u->faces : nullptr);
  if (rc != NX_OK) return rc;
  rc = run_pipeline(u, coarse_iters, lambda_min, s);
  if (rc != NX_OK) return rc;
  return nx::prolong_line_search_padded(sim, factor, u->W, u->H, u->out, u->mc, out, apply);
}

extern "C" {

int nx_learned_correct(nx_sim* sim, nx_unet* u, int factor, int coarse_iters,
                       float lambda_min, double* out) {
  return learned_correct_impl(sim, u, factor, coarse_iters, lambda_min, out, true);
}

int nx_learned_cycles(nx_sim* sim, nx_unet* u, int factor, int coarse_iters,
                      float lambda_min, int smooth_iters, float fine_lambda_min,
                      const int* tile, double target, double stall, int max_cycles,
                      double* records, int records_capacity, int* result) {
  // every argument is checked before anything is enqueued: a caller's
  // record buffer is written only within the capacity it states, and the
  // smoothing tile and interval are validated here rather than by the
  // first smoothing call (which runs after a correction was enqueued)
  if (!sim || !u || !result || factor < 1 || coarse_iters < 0 || smooth_iters < 1 ||
      max_cycles < 0 || !(stall > 0.0) || !(fine_lambda_min < 8.0f))
    return fail(NX_ERR_ARG, "bad learned cycle arguments");
  if (records && (records_capacity < 0 || (long long)records_capacity < 3LL * max_cycles))
    return fail(NX_ERR_ARG, "records_capacity must be at least 3 * max_cycles");
  // with Chebyshev smoothing the correction's step is taken by the first
  // smoothing launch as it loads the pressure (one pass less)
  const bool cheb = fine_lambda_min > 0.0f;
  if (cheb && tile && !nx::cheb_cfg_supported(nx::JacobiCfg{tile[0], tile[1], tile[2], tile[3]}))
    return fail(NX_ERR_ARG, "unsupported Chebyshev tile");
  int n = 0, kept = 0, converged = 0, abandoned = 0;
  double before = -1.0;
  for (int c = 0; c < max_cycles; ++c) {
    int rc = learned_correct_impl(sim, u, factor, coarse_iters, lambda_min, nullptr, !cheb);
    if (rc != NX_OK) return rc;
    double out[2] = {0.0, 0.0};
    rc = cheb ? nx::smooth_chebyshev_pending(sim, smooth_iters, fine_lambda_min, tile, out)
              : nx_smooth(sim, smooth_iters - 1, 0.5f, out);
    if (rc != NX_OK) return rc;
    ++n;
    const double rel = out[0], a = out[1];
    if (records) {
      nx_stats st{};
      rc = nx_get_stats(sim, &st);
      if (rc != NX_OK) return rc;
      records[3 * c] = a;
      records[3 * c + 1] = (double)st.sweeps;
      records[3 * c + 2] = rel;
    }
    if (rel <= target) {
      converged = 1;
      kept += a != 0.0;
      break;
    }
    if (a == 0.0 || (before >= 0.0 && rel > stall * before)) {
      abandoned = 1;
      break;
    }
    ++kept;
    before = rel;
  }
  result[0] = n;
  result[1] = kept;
  result[2] = converged;
  result[3] = abandoned;
  return NX_OK;
}

int nx_unet_set_levels(nx_unet* u, int mid_factor, int mid_iters, float mid_lambda_min) {
  if (!u) return fail(NX_ERR_ARG, "null argument");
  if (mid_factor == 0) {
    u->mid_f = 0;
    return NX_OK;
  }
  if (mid_factor < 2 || mid_iters < 1 || !(mid_lambda_min > 0.0f && mid_lambda_min < 8.0f))
    return fail(NX_ERR_ARG, "middle level: factor >= 2, iterations >= 1, 0 < lambda_min < 8");
  u->mid_f = mid_factor;
  u->mid_iters = mid_iters;
  u->mid_lmin = mid_lambda_min;
  return NX_OK;
}

int nx_unet_set_bypass(nx_unet* u, int bypass) {
  if (!u) return fail(NX_ERR_ARG, "null argument");
  u->bypass = bypass != 0;
  return NX_OK;
}

int nx_unet_coarse(nx_unet* u, float** rc, uint8_t** mc, float** e) {
  if (!u) return fail(NX_ERR_ARG, "null argument");
  if (rc) *rc = u->rc;
  if (mc) *mc = u->mc;
  if (e) *e = u->out;
  return NX_OK;
}

}  // extern "C"
