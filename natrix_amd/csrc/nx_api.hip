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
#include "nx_kernels.h"
#include "nx_jacobi_tile.h"

using namespace nx;

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int nx::set_error(int code, const std::string& msg) { return fail(code, msg); }

#define NX_HIP(call)                                                     \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess)                                                \
      return fail(NX_ERR_HIP, std::string(#call) + ": " +                \
                                  hipGetErrorString(e_) + " (" __FILE__ \
                                  ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define NX_TRY(expr)        \
  do {                      \
    int rc_ = (expr);       \
    if (rc_ != NX_OK) return rc_; \
  } while (0)

struct TimedPair {
  int tag;
  hipEvent_t a, b;
};

struct nx_sim {
  int W = 0, H = 0;
  int device = 0;
  int storage = ST_F32;  // format of velocity, pressure, divergence, dye
  size_t es = 4;         // bytes per stored scalar
  float* p32 = nullptr;  // float32 corrected pressure (reduced storage)
  hipStream_t stream = nullptr;
  float2* vel[2] = {nullptr, nullptr};
  float* p[2] = {nullptr, nullptr};
  float* div = nullptr;
  float* curl = nullptr;
  uint8_t* mask[2] = {nullptr, nullptr};
  float* dye[2] = {nullptr, nullptr};
  int dyeW = 0, dyeH = 0;
  int vcur = 0, pcur = 0, mcur = 0, dcur = 0;
  uint8_t* map_staging = nullptr;

  double* red5 = nullptr;  // residual_blocks * 5
  double* red2 = nullptr;  // jacobi blocks * 2
  size_t red2_cap = 0;
  double* sums = nullptr;  // 8 doubles
  int* state = nullptr;    // 4 ints
  int* host_state = nullptr;  // pinned, 2 slots x 4 ints
  double* host_out = nullptr;  // pinned, 8 doubles: per-call scalar read-backs
  hipEvent_t poll_ev[2] = {nullptr, nullptr};

  nx_params prm;
  nx_stats stats;

  bool timing = false;
  std::vector<TimedPair> timed;
  std::vector<hipEvent_t> free_events;

  // iterative refinement: pressure = ph + pl (float32 pair), original rhs
  float* ph = nullptr;
  float* pl = nullptr;
  float* b0 = nullptr;
  float* dcorr = nullptr;   // prolonged coarse correction (line search)
  float* rrow = nullptr;    // restriction row records (2 x H x Wc)
  float* cws = nullptr;     // banded coarse smoother: x, r, d twice
  float* fcheb = nullptr;   // fine Chebyshev smoother: r, d twice (4 W H)
  double* cheb_red = nullptr;  // its residual records, 5 per block
  size_t cheb_red_cap = 0;
  size_t cws_cap = 0;
  size_t rrow_cap = 0;
  double* line = nullptr;   // {alpha, d.r, d.Ld, 0}
  // fine Chebyshev smoother's launch shape when the caller passes no tile
  // (nx_set_smoother_tile): tile kernel when seg == 0, streaming otherwise
  JacobiCfg cheb_cfg{8, 4, 8, 2, 0};
  bool refining = false;
  // step phases (nx_step_begin .. nx_step_finish)
  bool in_step = false;
  float phase_dt = 0.0f;
  int step_sweeps = 0;
  // graphs keyed by the parities of the velocity, mask and dye buffers
  bool graph_on = false;
  hipGraphExec_t graph[8] = {};
  float graph_dt[8] = {};
  nx_params graph_prm[8];
};

// ------------------------------------------------------------- helpers
static bool params_equal(const nx_params& a, const nx_params& b) {
  return memcmp(&a, &b, sizeof(nx_params)) == 0;
}

static int get_event(nx_sim* s, hipEvent_t* ev) {
  if (!s->free_events.empty()) {
    *ev = s->free_events.back();
    s->free_events.pop_back();
    return NX_OK;
  }
  NX_HIP(hipEventCreate(ev));
  return NX_OK;
}

struct TimeScope {
  nx_sim* s;
  int tag;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = NX_OK;
  TimeScope(nx_sim* s_, int tag_) : s(s_), tag(tag_) {
    if (!s->timing) return;
    rc = get_event(s, &a);
    if (rc == NX_OK) rc = get_event(s, &b);
    if (rc == NX_OK && hipEventRecord(a, s->stream) != hipSuccess)
      rc = fail(NX_ERR_HIP, "hipEventRecord failed");
  }
  int end() {
    if (!s->timing || rc != NX_OK) return rc;
    if (hipEventRecord(b, s->stream) != hipSuccess)
      return fail(NX_ERR_HIP, "hipEventRecord failed");
    s->timed.push_back({tag, a, b});
    return NX_OK;
  }
};

#define NX_LAUNCH_TIMED(sim, tag, launch)         \
  do {                                            \
    TimeScope ts_(sim, tag);                      \
    if (ts_.rc != NX_OK) return ts_.rc;           \
    NX_HIP(launch);                               \
    NX_TRY(ts_.end());                            \
  } while (0)

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
static int run_jacobi_tb(nx_sim* s, int sweeps, JacobiCfg c) {
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

// Row streaming for the residual-type kernels below: wave w of a 64 x
// RES_ROWS block walks rows y0 .. y0 + RES_ROWS/4 - 1 of column x = 64 bx +
// lane, holding the rows above and below in registers; the left and right
// neighbours come from the adjacent lanes by DPP, lanes 0 and 63 load theirs
// (x - 1, x + 1) and the grid's edges use the cell itself, as the clamped
// index does.  Lanes past the last column read column W - 1 and write
// nothing, so every lane runs the same code.
struct RowNbr {
  float l, r;
  int ml, mr;
};

__device__ __forceinline__ RowNbr row_nbr(const float* __restrict__ p,
                                          const uint8_t* __restrict__ mask, int row, int x,
                                          int W, int lane, float pc, int mcv) {
  RowNbr n;
  n.l = jt::lane_from_left(pc);
  n.r = jt::lane_from_right(pc);
  n.ml = jt::ilane_from_left(mcv);
  n.mr = jt::ilane_from_right(mcv);
  if (lane == 0) {
    n.l = x > 0 ? p[row + x - 1] : pc;
    n.ml = x > 0 ? mask[row + x - 1] : mcv;
  }
  if (lane == 63 || x + 1 >= W) {
    n.r = x + 1 < W ? p[row + x + 1] : pc;
    n.mr = x + 1 < W ? mask[row + x + 1] : mcv;
  }
  return n;
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
  const int y0 = blockIdx.y * RES_ROWS + wave * (RES_ROWS / 4);
  if (y0 >= H) return;  // wave-uniform
  const int y1 = min(y0 + RES_ROWS / 4, H);
  float pc = p[y0 * W + xc];
  int mcv = mask[y0 * W + xc];
  float pu = y0 > 0 ? p[(y0 - 1) * W + xc] : pc;
  int mu = y0 > 0 ? mask[(y0 - 1) * W + xc] : mcv;
  for (int y = y0; y < y1; ++y) {
    const int row = y * W;
    const bool dn = y + 1 < H;
    const float pd = dn ? p[row + W + xc] : pc;
    const int md = dn ? mask[row + W + xc] : mcv;
    const RowNbr n = row_nbr(p, mask, row, x, W, lane, pc, mcv);
    float v = 0.0f, c = 0.0f;
    if (x < W && mcv == 0) {
      const float xl = n.ml ? pc : n.l;
      const float xr = n.mr ? pc : n.r;
      const float yb = mu ? pc : pu;
      const float yt = md ? pc : pd;
      v = b[row + x] - ((((xl + xr) + yb) + yt) - 4.0f * pc);
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
    pu = pc;
    mu = mcv;
    pc = pd;
    mcv = md;
  }
}

// Coarse cells of a padded coarse grid (Wp x Hp beyond the Wv x Hv cells
// that cover the fine grid) are solid with a zero right-hand side.
__global__ void __launch_bounds__(256)
    k_restrict_cols(const float* __restrict__ rsum, const float* __restrict__ rcnt,
                    float* __restrict__ rc, uint8_t* __restrict__ mc, int H, int f,
                    int Wv, int Hv, int Wp, int Hp) {
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
  rc[Y * Wp + X] = cnt == 0.0f ? 0.0f : (float)(f * f) * (sum / cnt);
  mc[Y * Wp + X] = cnt == 0.0f ? 1 : 0;
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
  float v = 0.0f;
  if (mask[pos] == 0) {
    const float cx = ((float)x - 0.5f * (float)(f - 1)) / (float)f;
    const float cy = ((float)y - 0.5f * (float)(f - 1)) / (float)f;
    const float fx = fminf(fmaxf(cx, 0.0f), (float)(Wc - 1));
    const float fy = fminf(fmaxf(cy, 0.0f), (float)(Hc - 1));
    const int x0 = (int)fx, y0 = (int)fy;
    const int x1 = min(x0 + 1, Wc - 1), y1 = min(y0 + 1, Hc - 1);
    const float tx = fx - (float)x0, ty = fy - (float)y0;
    const int i00 = y0 * Wc + x0, i01 = y0 * Wc + x1, i10 = y1 * Wc + x0,
              i11 = y1 * Wc + x1;
    const float w00 = mc[i00] ? 0.0f : (1.0f - tx) * (1.0f - ty);
    const float w01 = mc[i01] ? 0.0f : tx * (1.0f - ty);
    const float w10 = mc[i10] ? 0.0f : (1.0f - tx) * ty;
    const float w11 = mc[i11] ? 0.0f : tx * ty;
    const float wsum = ((w00 + w01) + w10) + w11;
    if (wsum > 0.0f)
      v = (((w00 * ec[i00] + w01 * ec[i01]) + w10 * ec[i10]) + w11 * ec[i11]) / wsum;
  }
  if (ADD) {
    if (mask[pos] == 0) p[pos] = p[pos] + scale * v;
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

  // out = L v for this lane's cells (all threads of the block call it)
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
        out[i][c] = (((L + R) + U) + D) - 4.0f * ctr;
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
    t.lap(x, lv);
#pragma unroll
    for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        r[i][c] = lv[i][c] - r[i][c];
        d[i][c] = r[i][c] / theta;
      }
    for (int s = 0; s < iters; ++s) {
      t.lap(d, lv);
      const float rho1 = 1.0f / (2.0f * sigma - rho);
      const float cd = rho1 * rho, cr = 2.0f * rho1 / delta;
#pragma unroll
      for (int i = 0; i < CJ_ROWS; ++i)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          x[i][c] = x[i][c] + d[i][c];
          r[i][c] = r[i][c] + lv[i][c];  // q -= A d
          d[i][c] = cd * d[i][c] + cr * r[i][c];
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

// The Chebyshev smoothing of k_coarse_smooth<true> spread over row bands,
// one workgroup per band of CB_BAND rows: a launch runs CB_T iterations
// (plus, at the first launch, the initial residual) on its band extended by
// CB_T + 1 rows on each side, whose outer rows go stale one row per stencil,
// and stores the band's x, r, d.  Launch boundaries are the only global
// synchronisation.  Every cell sees the operations of the single-workgroup
// kernel in the same order; theta and the coefficients of the launch's
// iterations (cf) come from the host, computed by the same float32
// operations (IEEE single on both sides).
constexpr int CB_ROWS = 2, CB_EY = CB_ROWS * CJ_WAVES, CB_T = 8;
constexpr int CB_BAND = CB_EY - 2 * (CB_T + 1);

__global__ void __launch_bounds__(1024)
    k_coarse_cheb_band(const float* __restrict__ xin, const float* __restrict__ rin,
                       const float* __restrict__ din, float* __restrict__ xout,
                       float* __restrict__ rout, float* __restrict__ dout,
                       const float* __restrict__ rhs, const uint8_t* __restrict__ mc,
                       int Wc, int Hc, int s0, int T, float theta, ChebCoef cf) {
  __shared__ float top[2][CJ_WAVES][CJ_MAX], bot[2][CJ_WAVES][CJ_MAX];
  CoarseTile<CB_ROWS> t;
  t.top = top;
  t.bot = bot;
  t.lane = threadIdx.x & 63;
  t.wave = threadIdx.x >> 6;
  t.x0 = 2 * t.lane;
  t.buf = 0;
  const int band0 = blockIdx.x * CB_BAND;
  const int y0 = band0 - (CB_T + 1) + CB_ROWS * t.wave;
  const bool first = s0 == 0;
  auto solid = [&](int x, int y) {
    return x < 0 || y < 0 || x >= Wc || y >= Hc || mc[y * Wc + x] != 0;
  };
  float x[CB_ROWS][2], r[CB_ROWS][2], d[CB_ROWS][2], lv[CB_ROWS][2];
#pragma unroll
  for (int i = 0; i < CB_ROWS; ++i)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gx = t.x0 + c, gy = y0 + i;
      const bool in = gx < Wc && gy >= 0 && gy < Hc;
      const int pos = in ? gy * Wc + gx : 0;
      x[i][c] = in ? xin[pos] : 0.0f;
      r[i][c] = in ? (first ? rhs[pos] : rin[pos]) : 0.0f;
      d[i][c] = (in && !first) ? din[pos] : 0.0f;
      t.code[i][c] = (solid(gx - 1, gy) ? 1 : 0) | (solid(gx + 1, gy) ? 2 : 0) |
                     (solid(gx, gy - 1) ? 4 : 0) | (solid(gx, gy + 1) ? 8 : 0);
    }
  if (first) {
    t.lap(x, lv);
#pragma unroll
    for (int i = 0; i < CB_ROWS; ++i)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        r[i][c] = lv[i][c] - r[i][c];
        d[i][c] = r[i][c] / theta;
      }
  }
  for (int it = 0; it < T; ++it) {
    t.lap(d, lv);
    const float cd = cf.cd[it], cr = cf.cr[it];
#pragma unroll
    for (int i = 0; i < CB_ROWS; ++i)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        x[i][c] = x[i][c] + d[i][c];
        r[i][c] = r[i][c] + lv[i][c];
        d[i][c] = cd * d[i][c] + cr * r[i][c];
      }
  }
#pragma unroll
  for (int i = 0; i < CB_ROWS; ++i)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int gx = t.x0 + c, gy = y0 + i;
      if (gx < Wc && gy >= band0 && gy < band0 + CB_BAND && gy < Hc) {
        const int pos = gy * Wc + gx;
        xout[pos] = x[i][c];
        rout[pos] = r[i][c];
        dout[pos] = d[i][c];
      }
    }
}

namespace nx {
int launch_coarse_smooth(hipStream_t st, int Wc, int Hc, float* e, const float* rc,
                         const uint8_t* mc, int iterations, float lambda_min, float* ws) {
  if (iterations == 0) return NX_OK;
  if (lambda_min > 0.0f && iterations > CB_T) {
    // banded: x, r, d ping-pong between the two halves of ws (6 Wc Hc
    // floats); the last launch writes x to e (read only by the first)
    if (!ws) return set_error(NX_ERR_ARG, "coarse smoother workspace missing");
    const size_t n = (size_t)Wc * Hc;
    float* set[2][3] = {{ws, ws + n, ws + 2 * n}, {ws + 3 * n, ws + 4 * n, ws + 5 * n}};
    const int nb = (Hc + CB_BAND - 1) / CB_BAND;
    ChebRecurrence rec(lambda_min);
    int done = 0, cur = 0;
    while (done < iterations) {
      const int T = iterations - done < CB_T ? iterations - done : CB_T;
      const bool last = done + T >= iterations;
      const float* xin = done == 0 ? e : set[cur][0];
      float** out = set[done == 0 ? 0 : 1 - cur];
      ChebCoef cf{};
      rec.next(0, T, cf);
      k_coarse_cheb_band<<<nb, 1024, 0, st>>>(xin, set[cur][1], set[cur][2],
                                              last ? e : out[0], out[1], out[2], rc, mc,
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
  // rows stream as in k_restrict_rows_seg; a solid neighbour is left out
  // (its difference to the centre is 0), a clamped one is the cell itself
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int x = blockIdx.x * 64 + lane, xc = min(x, W - 1);
  const int y0 = blockIdx.y * RES_ROWS + wave * (RES_ROWS / 4);
  const int y1 = min(y0 + RES_ROWS / 4, H);
  double v0 = 0.0, v1 = 0.0;
  if (y0 < H) {
    float pc = p[y0 * W + xc], dc = d[y0 * W + xc];
    int mcv = mask[y0 * W + xc];
    float pu = y0 > 0 ? p[(y0 - 1) * W + xc] : pc;
    float du = y0 > 0 ? d[(y0 - 1) * W + xc] : dc;
    int mu = y0 > 0 ? mask[(y0 - 1) * W + xc] : mcv;
    for (int y = y0; y < y1; ++y) {
      const int row = y * W;
      const bool dn = y + 1 < H;
      const float pd = dn ? p[row + W + xc] : pc;
      const float dd = dn ? d[row + W + xc] : dc;
      const int md = dn ? mask[row + W + xc] : mcv;
      const RowNbr np = row_nbr(p, mask, row, x, W, lane, pc, mcv);
      const RowNbr nd = row_nbr(d, mask, row, x, W, lane, dc, mcv);
      if (x < W && mcv == 0) {
        const double pcd = (double)pc, dcd = (double)dc;
        const float pv[4] = {np.l, np.r, pu, pd}, dv[4] = {nd.l, nd.r, du, dd};
        const int sv[4] = {np.ml, np.mr, mu, md};
        double lp = 0.0, ld = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (sv[j]) continue;
          lp += (double)pv[j] - pcd;
          ld += (double)dv[j] - dcd;
        }
        v0 += dcd * ((double)b[row + x] - lp);
        v1 += dcd * ld;
      }
      pu = pc;
      du = dc;
      mu = mcv;
      pc = pd;
      dc = dd;
      mcv = md;
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
  // rows stream as in k_restrict_rows_seg, with the pair (ph, pl) per cell
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int x = blockIdx.x * 64 + lane, xc = min(x, W - 1);
  const int y0 = blockIdx.y * RES_ROWS + wave * (RES_ROWS / 4);
  const int y1 = min(y0 + RES_ROWS / 4, H);
  double v[5] = {0, 0, 0, 0, 0};
  if (y0 < H) {
    float hc = ph[y0 * W + xc], lc = pl[y0 * W + xc];
    int mcv = mask[y0 * W + xc];
    float hu = y0 > 0 ? ph[(y0 - 1) * W + xc] : hc;
    float lu = y0 > 0 ? pl[(y0 - 1) * W + xc] : lc;
    int mu = y0 > 0 ? mask[(y0 - 1) * W + xc] : mcv;
    for (int y = y0; y < y1; ++y) {
      const int row = y * W;
      const bool dn = y + 1 < H;
      const float hd = dn ? ph[row + W + xc] : hc;
      const float ld = dn ? pl[row + W + xc] : lc;
      const int md = dn ? mask[row + W + xc] : mcv;
      const RowNbr nh = row_nbr(ph, mask, row, x, W, lane, hc, mcv);
      const RowNbr nl = row_nbr(pl, mask, row, x, W, lane, lc, mcv);
      if (x < W) {
        const int pos = row + x;
        const double pc = (double)hc + (double)lc;
        const double x1 = nh.ml ? pc : (double)nh.l + (double)nl.l;
        const double x2 = nh.mr ? pc : (double)nh.r + (double)nl.r;
        const double y1v = mu ? pc : (double)hu + (double)lu;
        const double y2v = md ? pc : (double)hd + (double)ld;
        const double bd = (double)b0[pos];
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
      hu = hc;
      lu = lc;
      mu = mcv;
      hc = hd;
      lc = ld;
      mcv = md;
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
                       nullptr, 0, 0, s->stream));
  s->host_out[1] = 0.0;
  NX_HIP(hipMemcpyAsync(s->host_out, s->sums + 5, sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  if (s->line)
    NX_HIP(hipMemcpyAsync(s->host_out + 1, s->line, sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
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
  NX_HIP(launch_reduce(s->cheb_red, last_blocks, 5, 0, s->sums, -1.0, nullptr, 0, 0,
                       s->stream));
  s->host_out[1] = 0.0;
  NX_HIP(hipMemcpyAsync(s->host_out, s->sums + 5, sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  if (s->line)
    NX_HIP(hipMemcpyAsync(s->host_out + 1, s->line, sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
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
static int restrict_impl(nx_sim* s, int factor, int Wp, int Hp, float* coarse_rhs,
                         uint8_t* coarse_mask) {
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
      rsum, rcnt, coarse_rhs, coarse_mask, s->H, factor, Wc, Hc, Wp, Hp);
  NX_HIP(hipGetLastError());
  return NX_OK;
}

int nx_restrict_residual(nx_sim* s, int factor, float* coarse_rhs,
                         uint8_t* coarse_mask) {
  if (!s || factor < 1) return fail(NX_ERR_ARG, "bad restriction arguments");
  return restrict_impl(s, factor, (s->W + factor - 1) / factor, (s->H + factor - 1) / factor,
                       coarse_rhs, coarse_mask);
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

int nx_coarse_smooth(nx_sim* s, int Wc, int Hc, float* e, const float* rc,
                     const uint8_t* coarse_mask, int iterations, float lambda_min) {
  if (!s || !e || !rc || !coarse_mask) return fail(NX_ERR_ARG, "null argument");
  if (Wc < 1 || Hc < 1 || Wc > CJ_MAX || Hc > CJ_ROWS * CJ_WAVES || iterations < 0)
    return fail(NX_ERR_ARG, "coarse grid must be at most 128 x 128, iterations >= 0");
  if (!(lambda_min < 8.0f)) return fail(NX_ERR_ARG, "lambda_min must be below 8");
  const size_t n = (size_t)Wc * Hc;
  if (lambda_min > 0.0f && iterations > CB_T && 6 * n > s->cws_cap) {
    if (s->cws) NX_HIP(hipFree(s->cws));
    s->cws = nullptr;
    NX_HIP(hipMalloc(&s->cws, 6 * n * sizeof(float)));
    s->cws_cap = 6 * n;
  }
  return nx::launch_coarse_smooth(s->stream, Wc, Hc, e, rc, coarse_mask, iterations,
                                  lambda_min, s->cws);
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
                       nullptr, 0, 0, s->stream));
  NX_HIP(hipMemcpyAsync(s->host_out, s->sums, 6 * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  NX_HIP(hipStreamSynchronize(s->stream));
  s->pcur = 0;
  *rel = s->host_out[5];
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

namespace nx {
int restrict_residual_padded(nx_sim* s, int factor, int Wp, int Hp, float* rc, uint8_t* mc) {
  return restrict_impl(s, factor, Wp, Hp, rc, mc);
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
