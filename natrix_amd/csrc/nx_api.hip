// Simulator object and C-ABI (include/natrix.h).
//
// Replaces the bgfx boundary of natrix/core/fluid_simulator.py: the
// reference keeps 7 GPU buffers (:357-368), binds them to shader slots
// (:338-355) and flips ping-pong pairs after each stencil pass (:444-474).
// Here one nx_sim owns the same fields as plain device arrays plus a HIP
// stream; each reference dispatch sequence becomes one function.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/natrix.h"
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"
#include "nx_sim.h"

using namespace nx;

static thread_local std::string g_last_error;

int nx::set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ------------------------------------------------------------- helpers
static bool params_equal(const nx_params& a, const nx_params& b) {
  return memcmp(&a, &b, sizeof(nx_params)) == 0;
}

static BcParams bc_of(const nx_params& p) {
  return BcParams{p.bc_mode, p.has_borders, p.inflow};
}

static BcParams bc_identity() { return BcParams{BC_REFERENCE, 0, 0.0f}; }

static int validate(const nx_sim* s, const nx_params& p) {
  // messages as natrix/core/fluid_simulator.py:62-111 (typos included)
  if (!(p.speed > 0)) return fail(NX_ERR_ARG, "'Speed' should be greater than zero");
  if (!(p.iterations > 0))
    return fail(NX_ERR_ARG, "'Iterations' should be grater than zero");
  if (!(p.dissipation > 0))
    return fail(NX_ERR_ARG, "'Dissipation' should be grater than zero");
  if (!(p.vorticity >= 0))
    return fail(NX_ERR_ARG, "'Vorticity' should be grater or equal than zero");
  if (!(p.viscosity >= 0))
    return fail(NX_ERR_ARG, "'Viscosity' should be greater or equal than zero");
  if (!(p.dye_speed > 0)) return fail(NX_ERR_ARG, "'Speed' should be greater than zero");
  if (!(p.dye_dissipation > 0))
    return fail(NX_ERR_ARG, "'Dissipation' should be grater than zero");
  if (p.bc_mode < 0 || p.bc_mode > 2) return fail(NX_ERR_ARG, "unknown bc_mode");
  if (p.bc_mode != BC_REFERENCE && s && (s->W < 3 || s->H < 3))
    return fail(NX_ERR_ARG, "free-slip and inflow walls need a grid of at least 3x3");
  JacobiCfg c{p.jacobi_k, p.jacobi_rows, p.jacobi_waves, p.jacobi_cols, p.jacobi_seg};
  if (p.jacobi_seg < 0 || !jacobi_cfg_supported(c))
    return fail(NX_ERR_ARG, "unsupported Jacobi tile (k, rows, waves, cols, seg)");
  if (!(p.tol >= 0)) return fail(NX_ERR_ARG, "tol must be >= 0");
  if (p.tol > 0 && !p.fused)
    return fail(NX_ERR_ARG, "the tolerance-driven solve needs the fused path");
  if (p.max_iters < 1) return fail(NX_ERR_ARG, "max_iters must be >= 1");
  if (p.correction_sweeps < 0 || p.correction_sweeps > 16)
    return fail(NX_ERR_ARG, "correction_sweeps must be in [0, 16]");
  if (s && s->storage != ST_F32) {
    if (!p.fused) return fail(NX_ERR_ARG, "reduced-precision storage runs the fused path only");
    if (p.jacobi_seg != 0)
      return fail(NX_ERR_ARG, "the streaming Jacobi kernel needs float32 storage");
    if (p.tol > 0)
      return fail(NX_ERR_ARG, "the tolerance-driven solve needs float32 storage");
    JacobiCfg cr{1 + p.correction_sweeps, p.jacobi_rows, p.jacobi_waves, p.jacobi_cols};
    if (!jacobi_cfg_supported(cr))
      return fail(NX_ERR_ARG, "correction_sweeps too large for the Jacobi tile");
  }
  if (p.check_every < 1) return fail(NX_ERR_ARG, "check_every must be >= 1");
  return NX_OK;
}

static int ensure_red2(nx_sim* s, JacobiCfg c) {
  const size_t need = (size_t)jacobi_launch_blocks(s->W, s->H, c) * 2;
  if (need <= s->red2_cap) return NX_OK;
  if (s->red2) NX_HIP(hipFree(s->red2));
  s->red2 = nullptr;
  NX_HIP(hipMalloc(&s->red2, need * sizeof(double)));
  s->red2_cap = need;
  return NX_OK;
}

static void invalidate_graphs(nx_sim* s) {
  for (int i = 0; i < 8; ++i) {
    if (s->graph[i]) (void)hipGraphExecDestroy(s->graph[i]);
    s->graph[i] = nullptr;
  }
}

// `sweeps` temporal-blocked sweeps starting from p[pcur]
int nx::run_jacobi_tb(nx_sim* s, int sweeps, JacobiCfg c) {
  while (sweeps > 0) {
    JacobiCfg cc = c;
    if (sweeps < cc.k) cc.k = sweeps;
    NX_LAUNCH_TIMED(s, NX_TIME_JACOBI,
                    launch_jacobi_tb(s->p[s->pcur], s->div, s->mask[s->mcur],
                                     s->p[1 - s->pcur], s->W, s->H, cc,
                                     nullptr, nullptr, s->stream));
    s->pcur ^= 1;
    sweeps -= cc.k;
  }
  return NX_OK;
}

static int state_reset(nx_sim* s) {
  // synchronous small copy: the tolerance-driven solve is never captured
  static const int zeros[4] = {0, 0, 0, 0};
  NX_HIP(hipMemcpyAsync(s->state, zeros, sizeof(zeros), hipMemcpyHostToDevice,
                        s->stream));
  NX_HIP(hipStreamSynchronize(s->stream));
  return NX_OK;
}

// Chunks of Jacobi launches, each chunk ending with a residual-accumulating
// launch and a reduction that raises the device `done` flag; later launches
// then exit at once.  The host polls the flag one chunk behind, so the GPU
// always has a chunk queued.  On return `sweeps` counts the sweeps applied
// to p[pcur].
static int iterate_chunks(nx_sim* s, JacobiCfg c, int& sweeps, bool& done) {
  const nx_params& P = s->prm;
  NX_TRY(state_reset(s));
  const int launches_per_chunk = (P.check_every + c.k - 1) / c.k;
  const int nblocks = jacobi_launch_blocks(s->W, s->H, c);
  int chunk = 0, polled = 0;
  int done_sweeps = -1, done_idx = -1;
  done = false;
  while (!done) {
    if (sweeps < P.max_iters) {
      for (int l = 0; l < launches_per_chunk && sweeps < P.max_iters; ++l) {
        JacobiCfg cc = c;
        if (P.max_iters - sweeps < cc.k) cc.k = P.max_iters - sweeps;
        const bool lastl = (l == launches_per_chunk - 1) ||
                           (sweeps + cc.k >= P.max_iters);
        // only a full-k launch has the tiling the partial buffer is sized for
        double* red = (lastl && cc.k == c.k) ? s->red2 : nullptr;
        NX_LAUNCH_TIMED(s, NX_TIME_JACOBI,
                        launch_jacobi_tb(s->p[s->pcur], s->div,
                                         s->mask[s->mcur], s->p[1 - s->pcur],
                                         s->W, s->H, cc, red, s->state,
                                         s->stream));
        s->pcur ^= 1;
        sweeps += cc.k;
        if (red) {
          // residual of iterate sweeps-1; this launch's output (iterate
          // `sweeps`) sits in p[pcur]
          NX_HIP(launch_reduce(s->red2, nblocks, 2, 0, s->sums, P.tol,
                               s->state, sweeps, s->pcur, s->stream));
        }
      }
      NX_HIP(hipMemcpyAsync(s->host_state + 4 * (chunk & 1), s->state,
                            4 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
      NX_HIP(hipEventRecord(s->poll_ev[chunk & 1], s->stream));
      ++chunk;
    }
    const bool at_cap = sweeps >= P.max_iters;
    if (chunk - polled >= 2 || (at_cap && polled < chunk)) {
      NX_HIP(hipEventSynchronize(s->poll_ev[polled & 1]));
      const int* hs = s->host_state + 4 * (polled & 1);
      if (hs[0]) {
        done = true;
        done_sweeps = hs[1];
        done_idx = hs[2];
      }
      ++polled;
    }
    if (at_cap && polled == chunk) break;
  }
  NX_HIP(hipStreamSynchronize(s->stream));
  if (done) {
    s->pcur = done_idx;
    sweeps = done_sweeps;
  }
  return NX_OK;
}

// exact FP64 relative residual of p[pcur] (synchronises)
static int exact_residual(nx_sim* s, double* rel) {
  NX_HIP(launch_sums(s->p[s->pcur], s->div, s->mask[s->mcur], s->W, s->H,
                     s->red5, s->stream));
  NX_HIP(launch_reduce(s->red5, residual_blocks(s->W, s->H), 5, 0, s->sums,
                       -1.0, nullptr, 0, 0, s->stream));
  NX_HIP(hipMemcpyAsync(s->host_out, s->sums, 6 * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  NX_HIP(hipStreamSynchronize(s->stream));
  *rel = s->host_out[5];
  return NX_OK;
}

// Tolerance-driven solve from p[pcur] (p1 of the fused kernel when
// `from_zero`, else a caller-provided guess).  The right-hand side is
// centred on its fluid mean first.  When the device flag reports
// convergence, the returned iterate's exact residual is checked once more
// (the flag measures the iterate one sweep earlier, and float32 round-off
// can move the last digits); iteration resumes if it is above tol.
static int solve_tolerance(nx_sim* s, int sweeps_done, bool from_zero) {
  const nx_params& P = s->prm;
  JacobiCfg c{P.jacobi_k, P.jacobi_rows, P.jacobi_waves, P.jacobi_cols, P.jacobi_seg};
  NX_TRY(ensure_red2(s, c));
  const int n = s->W * s->H;
  NX_HIP(launch_sums(nullptr, s->div, s->mask[s->mcur], s->W, s->H, s->red5,
                     s->stream));
  NX_HIP(launch_reduce(s->red5, residual_blocks(s->W, s->H), 5, 0, s->sums,
                       -1.0, nullptr, 0, 0, s->stream));
  NX_HIP(launch_center_rhs(s->div, from_zero ? s->p[s->pcur] : nullptr,
                           s->mask[s->mcur], n, s->sums, s->stream));
  int sweeps = sweeps_done;
  bool converged = false;
  double rel = 0.0;
  for (;;) {
    bool done = false;
    NX_TRY(iterate_chunks(s, c, sweeps, done));
    NX_TRY(exact_residual(s, &rel));
    if (rel <= P.tol) {
      converged = true;
      break;
    }
    if (!done || sweeps >= P.max_iters) break;
  }
  s->stats.sweeps = sweeps;
  s->stats.converged = converged ? 1 : 0;
  s->stats.residual = rel;
  s->stats.residual_sweeps = sweeps;
  return NX_OK;
}

static int step_reference_order(nx_sim* s, float dt) {
  // fluid_simulator.py:180-280, one kernel per shader
  const nx_params& P = s->prm;
  const int W = s->W, H = s->H;
  NX_HIP(launch_bc(s->vel[s->vcur], W, H, bc_of(P), s->stream));
  NX_LAUNCH_TIMED(s, NX_TIME_ADVECT,
                  launch_advect(s->vel[s->vcur], s->vel[1 - s->vcur],
                                s->mask[s->mcur], W, H, dt, P.speed,
                                P.dissipation, bc_identity(), s->stream));
  s->vcur ^= 1;
  NX_HIP(launch_curl(s->vel[s->vcur], s->curl, W, H, s->stream));
  NX_HIP(launch_confine(s->vel[s->vcur], s->curl, s->vel[1 - s->vcur], W, H,
                        dt, P.vorticity, s->stream));
  s->vcur ^= 1;
  if (P.viscosity > 0.0f) {
    const double a = 1.0 / (double)P.viscosity;
    const double rb = 1.0 / (4.0 + a);
    NX_HIP(launch_viscosity(s->vel[s->vcur], s->vel[1 - s->vcur], W, H,
                            (float)a, (float)rb, s->stream));
    s->vcur ^= 1;
  }
  NX_HIP(launch_divergence(s->vel[s->vcur], s->mask[s->mcur], s->div, W, H,
                           s->stream));
  NX_HIP(launch_fill(s->p[s->pcur], 0.0f, (size_t)W * H, s->stream));
  for (int i = 0; i < P.iterations; ++i) {
    NX_LAUNCH_TIMED(s, NX_TIME_JACOBI,
                    launch_jacobi_sweep(s->p[s->pcur], s->div,
                                        s->mask[s->mcur], s->p[1 - s->pcur],
                                        W, H, s->stream));
    s->pcur ^= 1;
  }
  NX_LAUNCH_TIMED(s, NX_TIME_PROJECT,
                  launch_subtract_gradient(s->vel[s->vcur], s->p[s->pcur],
                                           s->mask[s->mcur],
                                           s->vel[1 - s->vcur], nullptr, W, H,
                                           s->stream));
  s->vcur ^= 1;
  NX_HIP(launch_mask_keep(s->mask[s->mcur], MASK_PERSIST, (size_t)W * H,
                          s->stream));
  s->stats.sweeps = P.iterations;
  s->stats.converged = 0;
  if (s->dye[0] && P.dye_in_step) {
    NX_LAUNCH_TIMED(s, NX_TIME_DYE,
                    launch_advect_dye(s->dye[s->dcur], s->dye[1 - s->dcur],
                                      s->vel[s->vcur], s->mask[s->mcur],
                                      s->dyeW, s->dyeH, W, H, dt, P.dye_speed,
                                      P.dye_dissipation, s->stream));
    s->dcur ^= 1;
  }
  return NX_OK;
}

// Fused step: advect (walls folded into its reads), fused curl..p1,
// temporal-blocked Jacobi, projection + obstacle clear, dye advection.
// Reduced-precision step: the fused sequence with every field stored in
// bfloat16/float16, then the float32 residual-correction launch.
static int step_mixed_enqueue(nx_sim* s, float dt) {
  const nx_params& P = s->prm;
  const int W = s->W, H = s->H, st = s->storage;
  const int a = s->vcur;
  NX_LAUNCH_TIMED(s, NX_TIME_ADVECT,
                  launch_m_advect(st, s->vel[a], s->vel[1 - a], s->mask[s->mcur],
                                  W, H, dt, P.speed, P.dissipation, bc_of(P),
                                  s->stream));
  float alpha = 0.0f, rbeta = 0.0f;
  const int viscous = P.viscosity > 0.0f;
  if (viscous) {
    const double av = 1.0 / (double)P.viscosity;
    alpha = (float)av;
    rbeta = (float)(1.0 / (4.0 + av));
  }
  NX_LAUNCH_TIMED(s, NX_TIME_VORT_DIV,
                  launch_m_fused(st, s->vel[1 - a], s->mask[s->mcur], s->vel[a],
                                 s->div, s->p[0], W, H, dt, P.vorticity, viscous,
                                 alpha, rbeta, s->stream));
  s->pcur = 0;
  JacobiCfg c{P.jacobi_k, P.jacobi_rows, P.jacobi_waves, P.jacobi_cols, P.jacobi_seg};
  int sweeps = P.iterations - 1;
  while (sweeps > 0) {
    JacobiCfg cc = c;
    if (sweeps < cc.k) cc.k = sweeps;
    NX_LAUNCH_TIMED(s, NX_TIME_JACOBI,
                    launch_m_jacobi(st, s->p[s->pcur], s->div, s->mask[s->mcur],
                                    s->p[1 - s->pcur], nullptr, W, H, cc, false,
                                    s->stream));
    s->pcur ^= 1;
    sweeps -= cc.k;
  }
  // float32 residual correction: one residual pass + correction_sweeps
  // sweeps of A e = r, written as p + e (with 0 sweeps: the widened p)
  JacobiCfg cr = c;
  cr.k = 1 + P.correction_sweeps;
  NX_LAUNCH_TIMED(s, NX_TIME_JACOBI,
                  launch_m_jacobi(st, s->p[s->pcur], s->div, s->mask[s->mcur],
                                  nullptr, s->p32, W, H, cr, true, s->stream));
  s->stats.sweeps = P.iterations;
  s->stats.converged = 0;
  NX_LAUNCH_TIMED(s, NX_TIME_PROJECT,
                  launch_m_subgrad(st, s->vel[a], s->p32, s->mask[s->mcur],
                                   s->vel[1 - a], s->mask[1 - s->mcur], W, H,
                                   s->stream));
  s->vcur = 1 - a;
  s->mcur ^= 1;
  if (s->dye[0] && P.dye_in_step) {
    NX_LAUNCH_TIMED(s, NX_TIME_DYE,
                    launch_m_advect_dye(st, s->dye[s->dcur], s->dye[1 - s->dcur],
                                        s->vel[s->vcur], s->mask[s->mcur],
                                        s->dyeW, s->dyeH, W, H, dt, P.dye_speed,
                                        P.dye_dissipation, s->stream));
    s->dcur ^= 1;
  }
  return NX_OK;
}

static int step_fused_enqueue(nx_sim* s, float dt, bool tolerance) {
  if (s->storage != ST_F32) return step_mixed_enqueue(s, dt);
  const nx_params& P = s->prm;
  const int W = s->W, H = s->H;
  const int a = s->vcur;
  NX_LAUNCH_TIMED(s, NX_TIME_ADVECT,
                  launch_advect(s->vel[a], s->vel[1 - a], s->mask[s->mcur], W,
                                H, dt, P.speed, P.dissipation, bc_of(P),
                                s->stream));
  float alpha = 0.0f, rbeta = 0.0f;
  const int viscous = P.viscosity > 0.0f;
  if (viscous) {
    const double av = 1.0 / (double)P.viscosity;
    alpha = (float)av;
    rbeta = (float)(1.0 / (4.0 + av));
  }
  NX_LAUNCH_TIMED(s, NX_TIME_VORT_DIV,
                  launch_fused_vort_div(s->vel[1 - a], s->mask[s->mcur],
                                        s->vel[a], s->div, s->p[0], W, H, dt,
                                        P.vorticity, viscous, alpha, rbeta,
                                        s->stream));
  s->pcur = 0;
  if (tolerance) {
    NX_TRY(solve_tolerance(s, 1, true));
  } else {
    JacobiCfg c{P.jacobi_k, P.jacobi_rows, P.jacobi_waves, P.jacobi_cols, P.jacobi_seg};
    NX_TRY(run_jacobi_tb(s, P.iterations - 1, c));
    s->stats.sweeps = P.iterations;
    s->stats.converged = 0;
  }
  NX_LAUNCH_TIMED(s, NX_TIME_PROJECT,
                  launch_subtract_gradient(s->vel[a], s->p[s->pcur],
                                           s->mask[s->mcur], s->vel[1 - a],
                                           s->mask[1 - s->mcur], W, H,
                                           s->stream));
  s->vcur = 1 - a;
  s->mcur ^= 1;
  if (s->dye[0] && P.dye_in_step) {
    NX_LAUNCH_TIMED(s, NX_TIME_DYE,
                    launch_advect_dye(s->dye[s->dcur], s->dye[1 - s->dcur],
                                      s->vel[s->vcur], s->mask[s->mcur],
                                      s->dyeW, s->dyeH, W, H, dt, P.dye_speed,
                                      P.dye_dissipation, s->stream));
    s->dcur ^= 1;
  }
  return NX_OK;
}

// Dye as an RGBA8 image, demo/shaders/demo.ComputeShader.comp:9-19: every
// channel is the dye value stored to an rgba8 (unsigned normalised) image,
// i.e. clamp(v, 0, 1) * 255 rounded to nearest (ties up; a NaN stores 0).
template <class S>
__global__ void k_dye_rgba8(const typename S::s1* __restrict__ q, uint32_t* __restrict__ out,
                            int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float v = fminf(fmaxf(S::ld(q, i), 0.0f), 1.0f);
  const uint32_t c = (uint32_t)floorf(v * 255.0f + 0.5f);
  out[i] = c | (c << 8) | (c << 16) | (c << 24);
}

// The demo's velocity arrows, demo/shaders/demo.QuiverFragmentShader.frag:
// 14-70, for every pixel of a Ww x Hw framebuffer (row y holds
// gl_FragCoord.y = y + 0.5), as an rgba8 image (1, 1, 1, alpha).  Each
// pixel samples the velocity at its arrow tile's centre exactly as _field
// does: fPos = (1 - pos / WindowSize) * VelocitySize, corners clamped to
// the grid, mix(a, b, (t, 0)) (which interpolates only the x component;
// the y component is the bottom-left corner's), scaled by -WindowSize /
// VelocitySize and by tile * 0.4; then _vector's distance to the arrow
// (shaft and two head strokes) and alpha = 1 - clamp(dist, 0, 1), rounded
// to 8 bits like the dye image.  float32 throughout, in the shader's order.
__device__ __forceinline__ float qmix(float a, float b, float t) {
  return a * (1.0f - t) + b * t;
}

__device__ __forceinline__ float qlen(float x, float y) { return sqrtf(x * x + y * y); }

__device__ __forceinline__ float qline(float px, float py, float p1x, float p1y, float p2x,
                                       float p2y) {
  const float cx = (p1x + p2x) * 0.5f, cy = (p1y + p2y) * 0.5f;
  const float len = qlen(p2x - p1x, p2y - p1y);
  const float dx = (p2x - p1x) / len, dy = (p2y - p1y) / len;
  const float rx = px - cx, ry = py - cy;
  const float d1 = fabsf(rx * dy + ry * (-dx));
  const float d2 = fabsf(rx * dx + ry * dy) - 0.5f * len;
  return fmaxf(d1, d2);
}

template <class S>
__global__ void k_quiver_rgba8(const typename S::s2* __restrict__ u, int Wv, int Hv, int Ww,
                               int Hw, float T, uint32_t* __restrict__ out) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= Ww || y >= Hw) return;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const float cx = (floorf(px / T) + 0.5f) * T, cy = (floorf(py / T) + 0.5f) * T;
  const float Vw = (float)Wv, Vh = (float)Hv, Fw = (float)Ww, Fh = (float)Hw;
  const float fx = (1.0f - cx / Fw) * Vw, fy = (1.0f - cy / Fh) * Vh;
  const int trx = (int)fminf(fmaxf(ceilf(fx), 0.0f), Vw - 1.0f);
  const int tr_y = (int)fminf(fmaxf(ceilf(fy), 0.0f), Vh - 1.0f);
  const int blx = (int)fminf(fmaxf(floorf(fx), 0.0f), Vw - 1.0f);
  const int bly = (int)fminf(fmaxf(floorf(fy), 0.0f), Vh - 1.0f);
  const float dx = fx - (float)blx, dy = fy - (float)bly;
  const float2 lt = S::ld2(u, tr_y * Wv + blx), rt = S::ld2(u, tr_y * Wv + trx);
  const float2 lb = S::ld2(u, bly * Wv + blx), rb = S::ld2(u, bly * Wv + trx);
  const float h1x = qmix(lt.x, rt.x, dx), h1y = qmix(lt.y, rt.y, 0.0f);
  const float h2x = qmix(lb.x, rb.x, dx), h2y = qmix(lb.y, rb.y, 0.0f);
  const float mx = qmix(h2x, h1x, dy), my = qmix(h2y, h1y, 0.0f);
  float vx = (((-1.0f * mx) * (Fw / Vw)) * T) * 0.4f;
  float vy = (((-1.0f * my) * (Fh / Vh)) * T) * 0.4f;
  const float rx = px - cx, ry = py - cy;
  float dist = 1.0f;
  const float mag = qlen(vx, vy);
  if (mag > 0.001f) {
    const float ux = vx / mag, uy = vy / mag;
    const float mc = fminf(fmaxf(mag, 0.0f), T * 0.5f);
    vx = ux * mc;
    vy = uy * mc;
    const float shaft = qline(rx, ry, vx, vy, -vx, -vy);
    const float head = fminf(qline(rx, ry, vx, vy, 0.4f * vx + 0.2f * (-vy), 0.4f * vy + 0.2f * vx),
                             qline(rx, ry, vx, vy, 0.4f * vx + 0.2f * vy, 0.4f * vy + 0.2f * (-vx)));
    dist = fminf(shaft, head);
  }
  const float alpha = 1.0f - fminf(fmaxf(dist, 0.0f), 1.0f);
  const uint32_t a8 = (uint32_t)floorf(alpha * 255.0f + 0.5f);
  out[y * Ww + x] = 0x00FFFFFFu | (a8 << 24);
}

// ------------------------------------------------------------- C-ABI
extern "C" {

const char* nx_last_error(void) { return g_last_error.c_str(); }

int nx_default_params(nx_params* out) {
  if (!out) return fail(NX_ERR_ARG, "null params");
  nx_params p;
  memset(&p, 0, sizeof(p));
  p.speed = 500.0f;  // fluid_simulator.py:28-35
  p.iterations = 50;
  p.dissipation = 1.0f;
  p.vorticity = 0.0f;
  p.viscosity = 0.1f;
  p.has_borders = 1;
  p.bc_mode = BC_REFERENCE;
  p.inflow = 1.0f;
  p.dye_speed = 500.0f;  // smooth_particles_area.py:23-24
  p.dye_dissipation = 1.0f;
  p.dye_in_step = 1;
  p.fused = 1;
  p.jacobi_k = 6;
  p.jacobi_rows = 4;
  p.jacobi_waves = 8;
  p.jacobi_cols = 2;
  p.tol = 0.0;
  p.max_iters = 1000000;
  p.check_every = 256;
  p.correction_sweeps = 8;
  p.jacobi_seg = 0;
  *out = p;
  return NX_OK;
}

int nx_create(int width, int height, nx_sim** out) {
  return nx_create_ex(width, height, NX_STORAGE_F32, out);
}

int nx_create_ex(int width, int height, int storage, nx_sim** out) {
  if (!out) return fail(NX_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (width < 1 || height < 1 || (long long)width * height > (1ll << 28))
    return fail(NX_ERR_ARG, "grid size out of range");
  if (storage != NX_STORAGE_F32 && storage != NX_STORAGE_BF16 &&
      storage != NX_STORAGE_F16)
    return fail(NX_ERR_ARG, "unknown storage format");
  nx_sim* s = new nx_sim();
  s->W = width;
  s->H = height;
  s->storage = storage;
  s->es = storage == NX_STORAGE_F32 ? 4 : 2;
  nx_default_params(&s->prm);
  memset(&s->stats, 0, sizeof(s->stats));
  const size_t n = (size_t)width * height;
  auto cleanup = [&](int rc) {
    nx_destroy(s);
    return rc;
  };
  if (hipGetDevice(&s->device) != hipSuccess)
    return cleanup(fail(NX_ERR_HIP, "no HIP device"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(NX_ERR_HIP, "hipStreamCreate failed"));
  bool ok = true;
  const size_t es = s->es;
  for (int i = 0; i < 2; ++i) {
    ok &= hipMalloc(&s->vel[i], n * 2 * es) == hipSuccess;
    ok &= hipMalloc(&s->p[i], n * es) == hipSuccess;
    ok &= hipMalloc(&s->mask[i], n) == hipSuccess;
  }
  ok &= hipMalloc(&s->div, n * es) == hipSuccess;
  ok &= hipMalloc(&s->curl, n * sizeof(float)) == hipSuccess;
  if (storage != NX_STORAGE_F32) ok &= hipMalloc(&s->p32, n * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&s->red5, (size_t)residual_blocks(width, height) * 5 *
                                sizeof(double)) == hipSuccess;
  ok &= hipMalloc(&s->sums, 8 * sizeof(double)) == hipSuccess;
  ok &= hipMalloc(&s->state, 4 * sizeof(int)) == hipSuccess;
  ok &= hipHostMalloc(&s->host_state, 8 * sizeof(int), 0) == hipSuccess;
  ok &= hipHostMalloc(&s->host_out, 8 * sizeof(double), 0) == hipSuccess;
  ok &= ok && hipHostGetDevicePointer(reinterpret_cast<void**>(&s->host_out_dev), s->host_out,
                                      0) == hipSuccess;
  ok &= hipEventCreateWithFlags(&s->poll_ev[0], hipEventDisableTiming) == hipSuccess;
  ok &= hipEventCreateWithFlags(&s->poll_ev[1], hipEventDisableTiming) == hipSuccess;
  if (!ok) return cleanup(fail(NX_ERR_HIP, "device allocation failed"));
  // zero state: fields start at rest, as freshly created bgfx buffers
  for (int i = 0; i < 2; ++i) {
    ok &= hipMemsetAsync(s->vel[i], 0, n * 2 * es, s->stream) == hipSuccess;
    ok &= hipMemsetAsync(s->p[i], 0, n * es, s->stream) == hipSuccess;
    ok &= hipMemsetAsync(s->mask[i], 0, n, s->stream) == hipSuccess;
  }
  ok &= hipMemsetAsync(s->div, 0, n * es, s->stream) == hipSuccess;
  if (s->p32) ok &= hipMemsetAsync(s->p32, 0, n * sizeof(float), s->stream) == hipSuccess;
  ok &= hipMemsetAsync(s->curl, 0, n * sizeof(float), s->stream) == hipSuccess;
  ok &= hipMemsetAsync(s->sums, 0, 8 * sizeof(double), s->stream) == hipSuccess;
  ok &= hipMemsetAsync(s->state, 0, 4 * sizeof(int), s->stream) == hipSuccess;
  ok &= hipStreamSynchronize(s->stream) == hipSuccess;
  if (!ok) return cleanup(fail(NX_ERR_HIP, "device initialisation failed"));
  *out = s;
  return NX_OK;
}

int nx_destroy(nx_sim* s) {
  if (!s) return NX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  invalidate_graphs(s);
  for (auto& t : s->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : s->free_events) (void)hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) {
    (void)hipFree(s->vel[i]);
    (void)hipFree(s->p[i]);
    (void)hipFree(s->mask[i]);
    (void)hipFree(s->dye[i]);
    if (s->poll_ev[i]) (void)hipEventDestroy(s->poll_ev[i]);
  }
  (void)hipFree(s->div);
  (void)hipFree(s->p32);
  (void)hipFree(s->ph);
  (void)hipFree(s->pl);
  (void)hipFree(s->b0);
  (void)hipFree(s->dcorr);
  (void)hipFree(s->rrow);
  (void)hipFree(s->cws);
  (void)hipFree(s->fcheb);
  (void)hipFree(s->cheb_red);
  (void)hipFree(s->line);
  (void)hipFree(s->curl);
  (void)hipFree(s->red5);
  (void)hipFree(s->red2);
  (void)hipFree(s->sums);
  (void)hipFree(s->state);
  (void)hipFree(s->map_staging);
  if (s->host_state) (void)hipHostFree(s->host_state);
  if (s->host_out) (void)hipHostFree(s->host_out);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return NX_OK;
}

int nx_get_size(const nx_sim* s, int* width, int* height) {
  if (!s || !width || !height) return fail(NX_ERR_ARG, "null argument");
  *width = s->W;
  *height = s->H;
  return NX_OK;
}

int nx_get_params(const nx_sim* s, nx_params* out) {
  if (!s || !out) return fail(NX_ERR_ARG, "null argument");
  *out = s->prm;
  return NX_OK;
}

int nx_set_params(nx_sim* s, const nx_params* prm) {
  if (!s || !prm) return fail(NX_ERR_ARG, "null argument");
  NX_TRY(validate(s, *prm));
  s->prm = *prm;
  return NX_OK;
}

int nx_check_params(const nx_params* prm) {
  if (!prm) return fail(NX_ERR_ARG, "null params");
  return validate(nullptr, *prm);
}

int nx_enable_dye(nx_sim* s, int width, int height) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (width < 1 || height < 1 || (long long)width * height > (1ll << 28))
    return fail(NX_ERR_ARG, "dye grid size out of range");
  const size_t n = (size_t)width * height;
  NX_HIP(hipStreamSynchronize(s->stream));
  invalidate_graphs(s);
  for (int i = 0; i < 2; ++i) {
    if (s->dye[i]) NX_HIP(hipFree(s->dye[i]));
    s->dye[i] = nullptr;
  }
  for (int i = 0; i < 2; ++i) {
    NX_HIP(hipMalloc(&s->dye[i], n * s->es));
    NX_HIP(hipMemsetAsync(s->dye[i], 0, n * s->es, s->stream));
  }
  NX_HIP(hipStreamSynchronize(s->stream));
  s->dyeW = width;
  s->dyeH = height;
  s->dcur = 0;
  return NX_OK;
}

int nx_add_velocity(nx_sim* s, float px, float py, float vx, float vy,
                    float radius) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_LAUNCH_TIMED(s, NX_TIME_SPLAT,
                  launch_add_velocity(s->vel[s->vcur], s->vel[1 - s->vcur],
                                      s->W, s->H, px, py, vx, vy, radius,
                                      s->stream));
  s->vcur ^= 1;  // _flip_velocity_buffer, fluid_simulator.py:131
  return NX_OK;
}

int nx_add_circle_obstacle(nx_sim* s, float px, float py, float radius,
                           int is_static) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  (void)is_static;  // both branches of AddCircleObstacle.comp:28-35 agree
  NX_HIP(launch_circle(s->mask[s->mcur], s->W, s->H, px, py, radius,
                       s->stream));
  return NX_OK;
}

int nx_add_triangle_obstacle(nx_sim* s, float x1, float y1, float x2,
                             float y2, float x3, float y3, int is_static) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  NX_HIP(launch_triangle(s->mask[s->mcur], s->W, s->H, x1, y1, x2, y2, x3, y3,
                         is_static ? 1 : 0, s->stream));
  return NX_OK;
}

int nx_dye_rgba8(nx_sim* s, void* out) {
  if (!s || !out) return fail(NX_ERR_ARG, "null argument");
  if (!s->dye[0]) return fail(NX_ERR_STATE, "dye is not enabled");
  const int n = s->dyeW * s->dyeH;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (s->storage == ST_F32)
    k_dye_rgba8<F32S><<<blocks, 256, 0, s->stream>>>((const float*)s->dye[s->dcur],
                                                      (uint32_t*)out, n);
  else if (s->storage == ST_BF16)
    k_dye_rgba8<BF16S><<<blocks, 256, 0, s->stream>>>(
        (const BF16S::s1*)s->dye[s->dcur], (uint32_t*)out, n);
  else
    k_dye_rgba8<F16S><<<blocks, 256, 0, s->stream>>>(
        (const F16S::s1*)s->dye[s->dcur], (uint32_t*)out, n);
  NX_HIP(hipGetLastError());
  return NX_OK;
}

int nx_quiver_rgba8(nx_sim* s, int window_w, int window_h, float arrow_tile, void* out) {
  if (!s || !out) return fail(NX_ERR_ARG, "null argument");
  if (window_w < 1 || window_h < 1 || !(arrow_tile > 0.0f) || !(arrow_tile < 1e30f))
    return fail(NX_ERR_ARG, "window size >= 1 and a finite arrow tile > 0 required");
  const dim3 grid((window_w + 63) / 64, (window_h + 3) / 4);
  const void* u = s->vel[s->vcur];
  if (s->storage == ST_F32)
    k_quiver_rgba8<F32S><<<grid, dim3(64, 4), 0, s->stream>>>(
        (const F32S::s2*)u, s->W, s->H, window_w, window_h, arrow_tile, (uint32_t*)out);
  else if (s->storage == ST_BF16)
    k_quiver_rgba8<BF16S><<<grid, dim3(64, 4), 0, s->stream>>>(
        (const BF16S::s2*)u, s->W, s->H, window_w, window_h, arrow_tile, (uint32_t*)out);
  else
    k_quiver_rgba8<F16S><<<grid, dim3(64, 4), 0, s->stream>>>(
        (const F16S::s2*)u, s->W, s->H, window_w, window_h, arrow_tile, (uint32_t*)out);
  NX_HIP(hipGetLastError());
  return NX_OK;
}

int nx_add_dye(nx_sim* s, float px, float py, float amount, float radius) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  if (!s->dye[0]) return fail(NX_ERR_STATE, "dye is not enabled");
  NX_LAUNCH_TIMED(s, NX_TIME_SPLAT,
                  launch_add_dye(s->dye[s->dcur], s->dye[1 - s->dcur], s->dyeW,
                                 s->dyeH, px, py, amount, radius, s->stream));
  s->dcur ^= 1;
  return NX_OK;
}

int nx_add_gaussian_splats(nx_sim* s, const float* splats, int n) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (n < 0 || (n > 0 && !splats)) return fail(NX_ERR_ARG, "bad splat list");
  if (n == 0) return NX_OK;
  if (s->dye[0] && (s->dyeW != s->W || s->dyeH != s->H)) {
    for (int i = 0; i < n; ++i)
      if (splats[6 * i + 5] != 0.0f)
        return fail(NX_ERR_STATE,
                    "Gaussian dye splats need the dye grid to match the "
                    "velocity grid");
  }
  std::vector<Splat> v(n);
  for (int i = 0; i < n; ++i) {
    const float* f = splats + 6 * i;
    if (!(f[4] > 0)) return fail(NX_ERR_ARG, "splat radius must be > 0");
    v[i] = Splat{f[0], f[1], f[2], f[3], f[4], f[5]};
  }
  if (s->storage != ST_F32) {
    NX_LAUNCH_TIMED(s, NX_TIME_SPLAT,
                    launch_m_splats(s->storage, s->vel[s->vcur],
                                    s->dye[0] ? (void*)s->dye[s->dcur] : nullptr,
                                    s->W, s->H, v.data(), n, s->stream));
    return NX_OK;
  }
  NX_LAUNCH_TIMED(s, NX_TIME_SPLAT,
                  launch_gaussian_splats(s->vel[s->vcur],
                                         s->dye[0] ? s->dye[s->dcur] : nullptr,
                                         s->W, s->H, v.data(), n, s->stream));
  return NX_OK;
}

int nx_set_obstacle_map(nx_sim* s, const uint8_t* map) {
  if (!s || !map) return fail(NX_ERR_ARG, "null argument");
  const size_t n = (size_t)s->W * s->H;
  // merge on the host: current mask with bit2 replaced by the map
  std::vector<uint8_t> m(n);
  NX_HIP(hipStreamSynchronize(s->stream));
  NX_HIP(hipMemcpy(m.data(), s->mask[s->mcur], n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i)
    m[i] = (uint8_t)((m[i] & ~MASK_PERSIST) | (map[i] ? MASK_PERSIST : 0));
  NX_HIP(hipMemcpy(s->mask[s->mcur], m.data(), n, hipMemcpyHostToDevice));
  return NX_OK;
}

int nx_update(nx_sim* s, float dt) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  NX_TRY(validate(s, s->prm));
  const nx_params& P = s->prm;
  if (!P.fused) return step_reference_order(s, dt);
  if (P.tol > 0) return step_fused_enqueue(s, dt, true);
  if (!s->graph_on || s->timing) return step_fused_enqueue(s, dt, false);
  // graph replay keyed by the buffer parities and by dt + parameters
  const bool dye_step = s->dye[0] && P.dye_in_step;
  const int par = s->vcur | (s->mcur << 1) | (dye_step ? (s->dcur << 2) : 0);
  if (!s->graph[par] || s->graph_dt[par] != dt ||
      !params_equal(s->graph_prm[par], P)) {
    if (s->graph[par]) NX_HIP(hipGraphExecDestroy(s->graph[par]));
    s->graph[par] = nullptr;
    hipGraph_t g;
    NX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int vc = s->vcur, mc = s->mcur, dc = s->dcur, pc = s->pcur;
    int rc = step_fused_enqueue(s, dt, false);
    hipError_t ec = hipStreamEndCapture(s->stream, &g);
    // capture executed nothing: restore the buffer indices
    s->vcur = vc;
    s->mcur = mc;
    s->dcur = dc;
    s->pcur = pc;
    if (rc != NX_OK) return rc;
    if (ec != hipSuccess) return fail(NX_ERR_HIP, "hipStreamEndCapture failed");
    hipError_t ei = hipGraphInstantiate(&s->graph[par], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      s->graph[par] = nullptr;
      return fail(NX_ERR_HIP, "hipGraphInstantiate failed");
    }
    s->graph_dt[par] = dt;
    s->graph_prm[par] = P;
  }
  NX_HIP(hipGraphLaunch(s->graph[par], s->stream));
  // replay the index bookkeeping of step_fused_enqueue
  const int sweeps_tb = P.iterations - 1;
  int pc = 0, rem = sweeps_tb;
  while (rem > 0) {
    rem -= (rem < P.jacobi_k ? rem : P.jacobi_k);
    pc ^= 1;
  }
  s->pcur = pc;
  s->vcur ^= 1;
  s->mcur ^= 1;
  if (dye_step) s->dcur ^= 1;
  s->stats.sweeps = P.iterations;
  s->stats.converged = 0;
  return NX_OK;
}

int nx_step_begin(nx_sim* s, float dt) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_TRY(validate(s, s->prm));
  const nx_params& P = s->prm;
  const int W = s->W, H = s->H;
  const int a = s->vcur;
  NX_LAUNCH_TIMED(s, NX_TIME_ADVECT,
                  launch_advect(s->vel[a], s->vel[1 - a], s->mask[s->mcur], W,
                                H, dt, P.speed, P.dissipation, bc_of(P),
                                s->stream));
  float alpha = 0.0f, rbeta = 0.0f;
  const int viscous = P.viscosity > 0.0f;
  if (viscous) {
    const double av = 1.0 / (double)P.viscosity;
    alpha = (float)av;
    rbeta = (float)(1.0 / (4.0 + av));
  }
  NX_LAUNCH_TIMED(s, NX_TIME_VORT_DIV,
                  launch_fused_vort_div(s->vel[1 - a], s->mask[s->mcur],
                                        s->vel[a], s->div, s->p[0], W, H, dt,
                                        P.vorticity, viscous, alpha, rbeta,
                                        s->stream));
  s->pcur = 0;
  NX_HIP(launch_sums(nullptr, s->div, s->mask[s->mcur], W, H, s->red5, s->stream));
  NX_HIP(launch_reduce(s->red5, residual_blocks(W, H), 5, 0, s->sums, -1.0,
                       nullptr, 0, 0, s->stream));
  NX_HIP(launch_center_rhs(s->div, s->p[0], s->mask[s->mcur], W * H, s->sums,
                           s->stream));
  s->phase_dt = dt;
  s->in_step = true;
  s->step_sweeps = 1;
  return NX_OK;
}

int nx_solve(nx_sim* s, double tol, int max_sweeps, nx_stats* out) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->in_step) return fail(NX_ERR_STATE, "nx_solve outside nx_step_begin/finish");
  if (!(tol > 0) || max_sweeps < 1) return fail(NX_ERR_ARG, "bad tol / max_sweeps");
  nx_params saved = s->prm;
  s->prm.tol = tol;
  s->prm.max_iters = max_sweeps;
  JacobiCfg c{s->prm.jacobi_k, s->prm.jacobi_rows, s->prm.jacobi_waves, s->prm.jacobi_cols,
              s->prm.jacobi_seg};
  int rc = ensure_red2(s, c);
  int sweeps = 0;
  bool converged = false;
  double rel = 0.0;
  while (rc == NX_OK) {
    bool done = false;
    rc = iterate_chunks(s, c, sweeps, done);
    if (rc != NX_OK) break;
    rc = exact_residual(s, &rel);
    if (rc != NX_OK) break;
    if (rel <= tol) {
      converged = true;
      break;
    }
    if (!done || sweeps >= max_sweeps) break;
  }
  s->prm = saved;
  if (rc != NX_OK) return rc;
  s->step_sweeps += sweeps;
  s->stats.sweeps = s->step_sweeps;
  s->stats.converged = converged ? 1 : 0;
  s->stats.residual = rel;
  s->stats.residual_sweeps = s->step_sweeps;
  if (out) *out = s->stats;
  return NX_OK;
}

int nx_step_finish(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (!s->in_step) return fail(NX_ERR_STATE, "nx_step_finish without nx_step_begin");
  if (s->refining) NX_TRY(nx_refine_end(s));
  const nx_params& P = s->prm;
  const int W = s->W, H = s->H;
  const int a = s->vcur;
  NX_LAUNCH_TIMED(s, NX_TIME_PROJECT,
                  launch_subtract_gradient(s->vel[a], s->p[s->pcur],
                                           s->mask[s->mcur], s->vel[1 - a],
                                           s->mask[1 - s->mcur], W, H,
                                           s->stream));
  s->vcur = 1 - a;
  s->mcur ^= 1;
  if (s->dye[0] && P.dye_in_step) {
    NX_LAUNCH_TIMED(s, NX_TIME_DYE,
                    launch_advect_dye(s->dye[s->dcur], s->dye[1 - s->dcur],
                                      s->vel[s->vcur], s->mask[s->mcur],
                                      s->dyeW, s->dyeH, W, H, s->phase_dt,
                                      P.dye_speed, P.dye_dissipation, s->stream));
    s->dcur ^= 1;
  }
  s->in_step = false;
  return NX_OK;
}

int nx_set_graph(nx_sim* s, int on) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  NX_HIP(hipStreamSynchronize(s->stream));
  s->graph_on = on != 0;
  if (!s->graph_on) invalidate_graphs(s);
  return NX_OK;
}

// ---- per-stage entry points
int nx_stage_boundaries(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_TRY(validate(s, s->prm));
  NX_HIP(launch_bc(s->vel[s->vcur], s->W, s->H, bc_of(s->prm), s->stream));
  return NX_OK;
}

int nx_stage_advect(nx_sim* s, float dt) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_advect(s->vel[s->vcur], s->vel[1 - s->vcur], s->mask[s->mcur],
                       s->W, s->H, dt, s->prm.speed, s->prm.dissipation,
                       bc_identity(), s->stream));
  s->vcur ^= 1;
  return NX_OK;
}

int nx_stage_curl(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_curl(s->vel[s->vcur], s->curl, s->W, s->H, s->stream));
  return NX_OK;
}

int nx_stage_confine(nx_sim* s, float dt) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_confine(s->vel[s->vcur], s->curl, s->vel[1 - s->vcur], s->W,
                        s->H, dt, s->prm.vorticity, s->stream));
  s->vcur ^= 1;
  return NX_OK;
}

int nx_stage_viscosity(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  if (!(s->prm.viscosity > 0.0f))
    return fail(NX_ERR_STATE, "viscosity is 0: the reference skips this pass");
  const double a = 1.0 / (double)s->prm.viscosity;
  const double rb = 1.0 / (4.0 + a);
  NX_HIP(launch_viscosity(s->vel[s->vcur], s->vel[1 - s->vcur], s->W, s->H,
                          (float)a, (float)rb, s->stream));
  s->vcur ^= 1;
  return NX_OK;
}

int nx_stage_divergence(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_divergence(s->vel[s->vcur], s->mask[s->mcur], s->div, s->W,
                           s->H, s->stream));
  return NX_OK;
}

int nx_stage_clear_pressure(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_fill(s->p[s->pcur], 0.0f, (size_t)s->W * s->H, s->stream));
  return NX_OK;
}

int nx_stage_jacobi(nx_sim* s, int sweeps) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  if (sweeps < 0) return fail(NX_ERR_ARG, "sweeps must be >= 0");
  for (int i = 0; i < sweeps; ++i) {
    NX_LAUNCH_TIMED(s, NX_TIME_JACOBI,
                    launch_jacobi_sweep(s->p[s->pcur], s->div,
                                        s->mask[s->mcur], s->p[1 - s->pcur],
                                        s->W, s->H, s->stream));
    s->pcur ^= 1;
  }
  return NX_OK;
}

int nx_stage_subtract_gradient(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_subtract_gradient(s->vel[s->vcur], s->p[s->pcur],
                                  s->mask[s->mcur], s->vel[1 - s->vcur],
                                  nullptr, s->W, s->H, s->stream));
  s->vcur ^= 1;
  return NX_OK;
}

int nx_stage_clear_obstacles(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  NX_HIP(launch_mask_keep(s->mask[s->mcur], MASK_PERSIST, (size_t)s->W * s->H,
                          s->stream));
  return NX_OK;
}

int nx_stage_advect_dye(nx_sim* s, float dt) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  if (!s->dye[0]) return fail(NX_ERR_STATE, "dye is not enabled");
  NX_HIP(launch_advect_dye(s->dye[s->dcur], s->dye[1 - s->dcur],
                           s->vel[s->vcur], s->mask[s->mcur], s->dyeW,
                           s->dyeH, s->W, s->H, dt, s->prm.dye_speed,
                           s->prm.dye_dissipation, s->stream));
  s->dcur ^= 1;
  return NX_OK;
}

int nx_stage_fused_vort_div(nx_sim* s, float dt) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  float alpha = 0.0f, rbeta = 0.0f;
  const int viscous = s->prm.viscosity > 0.0f;
  if (viscous) {
    const double av = 1.0 / (double)s->prm.viscosity;
    alpha = (float)av;
    rbeta = (float)(1.0 / (4.0 + av));
  }
  NX_HIP(launch_fused_vort_div(s->vel[s->vcur], s->mask[s->mcur],
                               s->vel[1 - s->vcur], s->div, s->p[1 - s->pcur],
                               s->W, s->H, dt, s->prm.vorticity, viscous,
                               alpha, rbeta, s->stream));
  s->vcur ^= 1;
  s->pcur ^= 1;
  return NX_OK;
}

int nx_jacobi_blocked(nx_sim* s, int sweeps, int k, int rows, int waves,
                      int cols, int seg) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  if (sweeps < 0) return fail(NX_ERR_ARG, "sweeps must be >= 0");
  JacobiCfg c{k, rows, waves, cols, seg};
  if (!jacobi_cfg_supported(c))
    return fail(NX_ERR_ARG, "unsupported Jacobi tile (k, rows, waves, cols)");
  return run_jacobi_tb(s, sweeps, c);
}

// ---- state access
static int field_info(nx_sim* s, int field, void** ptr, size_t* bytes) {
  const size_t n = (size_t)s->W * s->H;
  switch (field) {
    case NX_FIELD_VELOCITY:
      *ptr = s->vel[s->vcur];
      *bytes = n * 2 * s->es;
      return NX_OK;
    case NX_FIELD_PRESSURE:
      *ptr = s->p[s->pcur];
      *bytes = n * s->es;
      return NX_OK;
    case NX_FIELD_PRESSURE32:
      *ptr = s->p32 ? s->p32 : s->p[s->pcur];
      *bytes = n * sizeof(float);
      return NX_OK;
    case NX_FIELD_DIVERGENCE:
      *ptr = s->div;
      *bytes = n * s->es;
      return NX_OK;
    case NX_FIELD_CURL:
      *ptr = s->curl;
      *bytes = n * sizeof(float);
      return NX_OK;
    case NX_FIELD_MASK:
      *ptr = s->mask[s->mcur];
      *bytes = n;
      return NX_OK;
    case NX_FIELD_DYE:
      if (!s->dye[0]) return fail(NX_ERR_STATE, "dye is not enabled");
      *ptr = s->dye[s->dcur];
      *bytes = (size_t)s->dyeW * s->dyeH * s->es;
      return NX_OK;
  }
  return fail(NX_ERR_ARG, "unknown field id");
}

int nx_field_ptr(nx_sim* s, int field, void** dev_ptr, size_t* bytes) {
  if (!s || !dev_ptr || !bytes) return fail(NX_ERR_ARG, "null argument");
  return field_info(s, field, dev_ptr, bytes);
}

int nx_upload(nx_sim* s, int field, const void* src, size_t bytes) {
  if (!s || !src) return fail(NX_ERR_ARG, "null argument");
  void* ptr;
  size_t nb;
  NX_TRY(field_info(s, field, &ptr, &nb));
  if (bytes != nb) return fail(NX_ERR_ARG, "byte count does not match the field");
  NX_HIP(hipStreamSynchronize(s->stream));
  NX_HIP(hipMemcpyAsync(ptr, src, nb, hipMemcpyHostToDevice, s->stream));
  NX_HIP(hipStreamSynchronize(s->stream));
  return NX_OK;
}

int nx_download(nx_sim* s, int field, void* dst, size_t bytes) {
  if (!s || !dst) return fail(NX_ERR_ARG, "null argument");
  void* ptr;
  size_t nb;
  NX_TRY(field_info(s, field, &ptr, &nb));
  if (bytes != nb) return fail(NX_ERR_ARG, "byte count does not match the field");
  NX_HIP(hipStreamSynchronize(s->stream));
  NX_HIP(hipMemcpyAsync(dst, ptr, nb, hipMemcpyDeviceToHost, s->stream));
  NX_HIP(hipStreamSynchronize(s->stream));
  return NX_OK;
}

int nx_residual(nx_sim* s, double* out) {
  if (!s || !out) return fail(NX_ERR_ARG, "null argument");
  if (s->storage != ST_F32)
    return fail(NX_ERR_STATE, "this entry point needs float32 storage");
  NX_HIP(launch_sums(s->p[s->pcur], s->div, s->mask[s->mcur], s->W, s->H,
                     s->red5, s->stream));
  NX_HIP(launch_reduce(s->red5, residual_blocks(s->W, s->H), 5, 0, s->sums,
                       -1.0, nullptr, 0, 0, s->stream));
  NX_HIP(hipMemcpyAsync(out, s->sums, 6 * sizeof(double),
                        hipMemcpyDeviceToHost, s->stream));
  NX_HIP(hipStreamSynchronize(s->stream));
  return NX_OK;
}

int nx_get_stats(nx_sim* s, nx_stats* out) {
  if (!s || !out) return fail(NX_ERR_ARG, "null argument");
  NX_HIP(hipStreamSynchronize(s->stream));
  *out = s->stats;
  return NX_OK;
}

int nx_synchronize(nx_sim* s) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  NX_HIP(hipStreamSynchronize(s->stream));
  return NX_OK;
}

int nx_get_stream(nx_sim* s, void** stream) {
  if (!s || !stream) return fail(NX_ERR_ARG, "null argument");
  *stream = (void*)s->stream;
  return NX_OK;
}

int nx_set_timing(nx_sim* s, int on) {
  if (!s) return fail(NX_ERR_ARG, "null simulator");
  NX_HIP(hipStreamSynchronize(s->stream));
  s->timing = on != 0;
  return NX_OK;
}

int nx_collect_timing(nx_sim* s, double* ms, int* launches) {
  if (!s || !ms || !launches) return fail(NX_ERR_ARG, "null argument");
  NX_HIP(hipStreamSynchronize(s->stream));
  for (int i = 0; i < NX_TIME_COUNT; ++i) {
    ms[i] = 0.0;
    launches[i] = 0;
  }
  for (auto& t : s->timed) {
    float e = 0.0f;
    NX_HIP(hipEventElapsedTime(&e, t.a, t.b));
    if (t.tag >= 0 && t.tag < NX_TIME_COUNT) {
      ms[t.tag] += e;
      launches[t.tag] += 1;
    }
    s->free_events.push_back(t.a);
    s->free_events.push_back(t.b);
  }
  s->timed.clear();
  return NX_OK;
}

}  // extern "C"
