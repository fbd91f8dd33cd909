// Host-side launchers for the fluid-step kernels.  Every launcher enqueues
// on the given stream and returns hipGetLastError() of the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "nx_cells.h"

struct nx_sim;

namespace nx {

// record `msg` as the calling thread's nx_last_error() and return `code`
int set_error(int code, const std::string& msg);

// ---- per-stage kernels (one reference shader each) -------------------
hipError_t launch_bc(float2* u, int W, int H, BcParams bc, hipStream_t s);
hipError_t launch_advect(const float2* uin, float2* uout, const uint8_t* mask,
                         int W, int H, float dt, float speed, float diss,
                         BcParams bc, hipStream_t s);
hipError_t launch_curl(const float2* u, float* w, int W, int H, hipStream_t s);
hipError_t launch_confine(const float2* uin, const float* w, float2* uout,
                          int W, int H, float dt, float scale, hipStream_t s);
hipError_t launch_viscosity(const float2* uin, float2* uout, int W, int H,
                            float alpha, float rbeta, hipStream_t s);
hipError_t launch_divergence(const float2* u, const uint8_t* mask, float* div,
                             int W, int H, hipStream_t s);
hipError_t launch_jacobi_sweep(const float* pin, const float* b,
                               const uint8_t* mask, float* pout, int W, int H,
                               hipStream_t s);
hipError_t launch_subtract_gradient(const float2* uin, const float* p,
                                    const uint8_t* mask, float2* uout,
                                    uint8_t* mask_next, int W, int H,
                                    hipStream_t s);
hipError_t launch_fill(float* a, float v, size_t n, hipStream_t s);
hipError_t launch_mask_keep(uint8_t* m, uint8_t keep, size_t n, hipStream_t s);
hipError_t launch_add_velocity(const float2* uin, float2* uout, int W, int H,
                               float px, float py, float vx, float vy,
                               float r, hipStream_t s);
hipError_t launch_circle(uint8_t* mask, int W, int H, float px, float py,
                         float r, hipStream_t s);
hipError_t launch_triangle(uint8_t* mask, int W, int H, float x1, float y1,
                           float x2, float y2, float x3, float y3,
                           int is_static, hipStream_t s);
hipError_t launch_advect_dye(const float* qin, float* qout, const float2* u,
                             const uint8_t* mask, int Wp, int Hp, int Wv,
                             int Hv, float dt, float speed, float diss,
                             hipStream_t s);
hipError_t launch_add_dye(const float* qin, float* qout, int W, int H,
                          float px, float py, float val, float r,
                          hipStream_t s);
// `splats` is a host array (passed to the kernels by value)
hipError_t launch_gaussian_splats(float2* u, float* q, int W, int H,
                                  const Splat* splats, int n, hipStream_t s);

// ---- fused fast path ---------------------------------------------------
// curl + confinement (+ viscosity) + divergence + first Jacobi sweep from a
// zero guess: writes the projected-step velocity input, b and p1.
hipError_t launch_fused_vort_div(const float2* uin, const uint8_t* mask,
                                 float2* uout, float* div, float* p1, int W,
                                 int H, float dt, float scale, int viscous,
                                 float alpha, float rbeta, hipStream_t s);

// Jacobi sweeps: cfg.k consecutive sweeps per launch (temporal blocking),
// bitwise equal to k single sweeps.  If `red` is non-null the launch also
// writes FP64 residual partials of its LAST sweep's input iterate to
// red[2*block + {0,1}] (sum r, sum r^2 over fluid cells).  `done`
// (nullable) is a device flag: when non-zero the launch returns at once.
struct JacobiCfg {
  int k;        // sweeps per launch
  int rows;     // rows per wave
  int waves;    // waves per block
  int cols;     // columns per lane
  int seg = 0;  // > 0: streaming kernel, output rows per workgroup segment
};
int jacobi_launch_blocks(int W, int H, JacobiCfg cfg);
hipError_t launch_jacobi_tb(const float* pin, const float* b,
                            const uint8_t* mask, float* pout, int W, int H,
                            JacobiCfg cfg, double* red, const int* done,
                            hipStream_t s);
bool jacobi_cfg_supported(JacobiCfg cfg);

// Restriction of the residual onto, and interpolation with the energy line
// search from, a Wp x Hp coarse grid padded beyond ceil(W / f) x ceil(H / f)
// (padded coarse cells are solid, with a zero right-hand side).  With
// faces (2 Wp Hp floats): block sums and the face weights of the weighted
// coarse operator.
int restrict_residual_padded(::nx_sim* s, int factor, int Wp, int Hp, float* rc, uint8_t* mc,
                             float* faces = nullptr);
// Middle level of a three-level cycle: block sums of a restricted field
// onto a grid f times coarser (padded cells solid), the coarse faces of
// the fine mask at factor f on a Wp x Hp grid, and bilinear interpolation
// e = ec (no line search) onto a W x H grid with mask m.
int block_sum_restrict(hipStream_t st, const float* rm, const uint8_t* mm, int Wm, int Hm, int f,
                       float* rc, uint8_t* mc, int Wc, int Hc);
int coarse_faces_padded(::nx_sim* s, int f, int Wp, int Hp, float* faces);
int prolong_to(hipStream_t st, float* e, const uint8_t* m, int W, int H, int f, const float* ec,
               const uint8_t* mc, int Wc, int Hc);
int prolong_line_search_padded(::nx_sim* s, int factor, int Wp, int Hp, const float* e,
                               const uint8_t* mc, double* out, bool apply = true);
// nx_smooth_chebyshev taking first the line-search step that a
// prolong_line_search_padded(..., apply = false) left pending
int smooth_chebyshev_pending(::nx_sim* s, int iterations, float lambda_min, const int* tile,
                             double* out);
// Chebyshev smoothing of the fine system L x = b (mask != 0: solid) on the
// interval [lambda_min, 8] of -L.  One launch runs c.k stencil passes on
// a (64 c.cols) x (c.rows c.waves) register tile per block: with setup != 0
// the first pass computes r = L x - b, d = r / theta from x and b (rin, din
// unread), every other pass s is an iteration x += d, r += L d,
// d = cf.cd[s] d + cf.cr[s] r.  rout == nullptr: only x is written.  With
// red, the last pass instead evaluates the exact FP64 residual b - L x of
// the final iterate on the written interior and each block writes
// red[5 blk + {sum r, sum r^2, sum b, sum b^2, fluid cells}] (blocks:
// jacobi_launch_blocks of the launch's shape and passes).  With dapp (first
// launch), the loaded x first takes a pending line-search step: x += a dapp
// on fluid cells, a = (float)line[0], nothing when a == 0.
constexpr int CHEB_MAX_PASSES = 32;
struct ChebCoef {
  float cd[CHEB_MAX_PASSES];
  float cr[CHEB_MAX_PASSES];
};
// the scalar recurrence (float32, on the host): theta, delta, sigma of the
// interval and rho of the next iteration; next() fills passes [first,
// passes) of a launch's coefficients and advances rho
struct ChebRecurrence {
  float theta, delta, sigma, rho;
  explicit ChebRecurrence(float lambda_min);
  void next(int first, int passes, ChebCoef& cf);
};
bool cheb_cfg_supported(JacobiCfg c);
// Streaming form of the same passes (nx_stream.hip): a workgroup walks a
// (64 cols)-column strip over `seg` output rows in chunks of rows x waves
// rows, carrying the two rows above each chunk per pass in LDS (no vertical
// halo re-reads).  Bitwise equal to launch_cheb_tb.  Records (red) are per
// workgroup: stream_launch_blocks of the launch's shape.
constexpr int STREAM_KMAX = 16;
bool stream_cfg_supported(JacobiCfg c);
int stream_launch_blocks(int W, int H, JacobiCfg c);
hipError_t launch_jacobi_stream(const float* pin, const float* b, const uint8_t* mask,
                                float* pout, int W, int H, JacobiCfg c, double* red,
                                const int* done, hipStream_t st);
hipError_t launch_cheb_stream(hipStream_t st, JacobiCfg c, const float* xin, const float* rin,
                              const float* din, const float* b, const uint8_t* mask,
                              float* xout, float* rout, float* dout, int W, int H, int setup,
                              float theta, const ChebCoef& cf, double* red,
                              const float* dapp = nullptr, const double* line = nullptr);
hipError_t launch_cheb_tb(hipStream_t st, JacobiCfg c, const float* xin, const float* rin,
                          const float* din, const float* b, const uint8_t* mask, float* xout,
                          float* rout, float* dout, int W, int H, int setup, float theta,
                          const ChebCoef& cf, double* red, const float* dapp = nullptr,
                          const double* line = nullptr);

// U-Net convolutions on weights tiled as wt[tap][ci / 16][co][16] (one
// contiguous block per K step and output-channel group); the same results
// as nx_conv3x3_fwd / nx_conv3x3_fwd_split with n = 1.
int conv3x3_tile_weights(const void* w, void* wt, int cin, int cout, hipStream_t s);
int head_fwd_scaled(const void* x, const float* w, const float* bias, float* out,
                    long long pixels, int c, const float* scal, hipStream_t s);
// With x2: the input is concat(upsample2(x) (c1 channels, h/2 x wd/2), x2
// (cin - c1 channels, h x wd)), read in the fragment loads (decoder layers).
// With ypool: also the 2 x 2 average pool of y (nx_avgpool2's values).
int conv3x3_fwd_tiled(const void* x, const void* wt, const float* bias, void* y, float* work,
                      int splits, int h, int wd, int cin, int cout, int relu, hipStream_t s,
                      const void* x2 = nullptr, int c1 = 0, void* ypool = nullptr);

// Coarse smoothing of L e = rc on a Wc x Hc (<= 128 x 128) grid: Jacobi
// sweeps (lambda_min <= 0) or Chebyshev on [lambda_min, 8]; more than 12
// Chebyshev iterations run on 2D tiles over many workgroups and need
// ws = 6 Wc Hc floats.  faces (2 Wc Hc floats, east then south weights):
// the face-weighted operator, always tiled, Chebyshev only.
int launch_coarse_smooth(hipStream_t st, int Wc, int Hc, float* e, const float* rc,
                         const uint8_t* mc, int iterations, float lambda_min, float* ws,
                         const float* faces = nullptr);

// Exact FP64 residual sums of `p` (of a zero guess when p == nullptr) per
// 64 x RES_ROWS block of cells (256 threads, RES_ROWS / 4 rows each):
// red[5*blk + {sum r, sum r^2, sum b, sum b^2, fluid count}].
constexpr int RES_ROWS = 32;
int residual_blocks(int W, int H);
hipError_t launch_sums(const float* p, const float* b, const uint8_t* mask,
                       int W, int H, double* red, hipStream_t s);
// Deterministic reduction of `nblocks` records of `width` doubles into
// sums[off..off+width), then sums[5] = relative residual from sums[0..4].
// With `state` (tolerance-driven solve): the first time rel <= tol it sets
// state[0]=1, state[1]=sweeps_now, state[2]=final_idx; state[3]=sweeps_now.
// With `mirror` (a device pointer to pinned host memory): mirror[0] = the
// relative residual and mirror[1] = extra[0] (0 without `extra`), written
// by the kernel, so the host reads them after synchronising the stream
// without a copy of its own.
hipError_t launch_reduce(const double* part, int nblocks, int width, int off,
                         double* sums, double tol, int* state, int sweeps_now,
                         int final_idx, hipStream_t s, double* mirror = nullptr,
                         const double* extra = nullptr);

// b -= fluid mean (sums[2]/sums[4]) and, if p != nullptr, p = first sweep
// of the centred system, on fluid cells only.
hipError_t launch_center_rhs(float* b, float* p, const uint8_t* mask, int n,
                             const double* sums, hipStream_t s);

// ---- reduced-precision storage (st = ST_BF16 or ST_F16) ------------------
hipError_t launch_m_advect(int st, const void* uin, void* uout, const uint8_t* mask,
                           int W, int H, float dt, float speed, float diss,
                           BcParams bc, hipStream_t s);
hipError_t launch_m_fused(int st, const void* uin, const uint8_t* mask, void* uout,
                          void* div, void* p1, int W, int H, float dt, float scale,
                          int viscous, float alpha, float rbeta, hipStream_t s);
// corr = false: k sweeps, p stored in `st`; corr = true: float32 residual
// correction (1 residual pass + k-1 sweeps), float32 p + e to pout32
hipError_t launch_m_jacobi(int st, const void* pin, const void* b, const uint8_t* mask,
                           void* pout, float* pout32, int W, int H, JacobiCfg c,
                           bool corr, hipStream_t s);
hipError_t launch_m_subgrad(int st, const void* uin, const float* p,
                            const uint8_t* mask, void* uout, uint8_t* mask_next,
                            int W, int H, hipStream_t s);
hipError_t launch_m_advect_dye(int st, const void* qin, void* qout, const void* u,
                               const uint8_t* mask, int Wp, int Hp, int Wv, int Hv,
                               float dt, float speed, float diss, hipStream_t s);
hipError_t launch_m_splats(int st, void* u, void* q, int W, int H, const Splat* sp,
                           int n, hipStream_t s);

}  // namespace nx
