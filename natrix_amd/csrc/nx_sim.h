// Private to the library: the simulator object behind the C-ABI's nx_sim
// handle and the error and timing helpers of the translation units that
// implement its entry points (nx_api.hip: creation, parameters, the step
// and its phases; nx_cycle.hip: the learned cycle's grid transfers,
// smoothers, line search and FP64 refinement).
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/natrix.h"
#include "nx_kernels.h"

// records msg as this thread's nx_last_error() text and returns code
inline int fail(int code, const std::string& msg) { return nx::set_error(code, msg); }

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
  int storage = nx::ST_F32;  // format of velocity, pressure, divergence, dye
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
  double* host_out_dev = nullptr;  // the same memory as the device addresses it
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
  float* cws = nullptr;     // tiled coarse smoother: x, r, d twice
  float* fcheb = nullptr;   // fine Chebyshev smoother: r, d twice (4 W H)
  double* cheb_red = nullptr;  // its residual records, 5 per block
  size_t cheb_red_cap = 0;
  size_t cws_cap = 0;
  size_t rrow_cap = 0;
  double* line = nullptr;   // {alpha, d.r, d.Ld, 0}
  // fine Chebyshev smoother's launch shape when the caller passes no tile
  // (nx_set_smoother_tile): tile kernel when seg == 0, streaming otherwise
  nx::JacobiCfg cheb_cfg{8, 4, 8, 2, 0};
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

inline int get_event(nx_sim* s, hipEvent_t* ev) {
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

namespace nx {
// `sweeps` Jacobi sweeps of the simulator's pressure with launch shape c
// (tile or streaming kernels, k sweeps per launch), on its stream
int run_jacobi_tb(::nx_sim* s, int sweeps, JacobiCfg c);
}  // namespace nx
