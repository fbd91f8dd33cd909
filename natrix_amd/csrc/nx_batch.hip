// Batched scenes: many independent grids of uneven sizes advanced together,
// one persistent launch per stage.
//
// All scenes' fields are packed back to back (scene s starts at cell offset
// off[s]).  Every stage kernel walks a virtual block space that concatenates
// the scenes' own tilings; a per-kernel prefix table (device memory) maps a
// virtual block to its scene by binary search, and the scene's parameter
// block (size, offsets, speed, dissipation, vorticity, walls, ...) is read
// from a device table.  The launched grid is the device's resident capacity
// for that kernel (CUs x occupancy), and each block grid-strides over the
// virtual blocks, so one launch covers every scene whatever their sizes.
// The per-cell and per-tile code is the same device code the single-grid
// kernels run, so each scene's result is bitwise the single-grid result.
//
// The virtual blocks go to the XCDs round-robin, not in the contiguous
// per-XCD runs of xcd_tile that the single-grid kernels use: tiles here
// differ in cost from scene to scene (size, obstacles, splats), and
// contiguous runs left some XCDs idle at the end of a launch (64 uneven
// scenes: 415 against 446 batched steps/s).  For the same reason the batched
// Jacobi tile keeps the occupancy its registers give.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/natrix.h"
#include "nx_fused_tile.h"
#include "nx_jacobi_tile.h"
#include "nx_kernels.h"

namespace nx {

struct SceneDev {
  int W, H;
  long long off;  // first cell of the scene in the packed arrays
  float speed, diss, vort, alpha, rbeta;
  int viscous;
  BcParams bc;
  float dye_speed, dye_diss;
  int splat_begin, splat_end;  // step splats of this scene
};

namespace {

__device__ __forceinline__ int find_scene(const int* __restrict__ prefix, int S,
                                          int vb) {
  int lo = 0, hi = S - 1;  // largest s with prefix[s] <= vb
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= vb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// pointwise kernels: 64x4 cell tiles per scene
#define NX_BATCH_CELL_LOOP(prefix, S, total)                                 \
  for (int vb = blockIdx.x; vb < (total); vb += gridDim.x) {                 \
    const int s = find_scene(prefix, S, vb);                                 \
    const SceneDev d = sc[s];                                                \
    const int lb = vb - prefix[s];                                           \
    const int nbx = (d.W + 63) / 64;                                         \
    const int x = (lb % nbx) * 64 + threadIdx.x;                             \
    const int y = (lb / nbx) * 4 + threadIdx.y;                              \
    if (x >= d.W || y >= d.H) continue;                                      \
    const long long pos = d.off + (long long)y * d.W + x;

__global__ void __launch_bounds__(256)
    kb_splats(const SceneDev* __restrict__ sc, int S,
              const int* __restrict__ prefix, int total, float2* __restrict__ u,
              float* __restrict__ q, const Splat* __restrict__ sp) {
  NX_BATCH_CELL_LOOP(prefix, S, total)
    if (d.splat_begin >= d.splat_end) continue;
    float2 v = u[pos];
    float dq = q ? q[pos] : 0.0f;
    for (int i = d.splat_begin; i < d.splat_end; ++i) {
      const Splat s = sp[i];
      const float w = gaussian_weight(s, x, y, d.W, d.H);
      v.x = v.x + s.ax * w;
      v.y = v.y + s.ay * w;
      dq = dq + s.dye * w;
    }
    u[pos] = v;
    if (q) q[pos] = dq;
  }
}

__global__ void __launch_bounds__(256)
    kb_advect(const SceneDev* __restrict__ sc, int S,
              const int* __restrict__ prefix, int total,
              const float2* __restrict__ uin, float2* __restrict__ uout,
              const uint8_t* __restrict__ mask, float dt) {
  NX_BATCH_CELL_LOOP(prefix, S, total)
    uout[pos] = advect_cell(uin + d.off, mask + d.off, x, y, d.W, d.H, dt,
                            d.speed, d.diss, d.bc);
  }
}

__global__ void __launch_bounds__(256)
    kb_subgrad(const SceneDev* __restrict__ sc, int S,
               const int* __restrict__ prefix, int total,
               const float2* __restrict__ uin, const float* __restrict__ p,
               const uint8_t* __restrict__ mask, float2* __restrict__ uout,
               uint8_t* __restrict__ mask_next) {
  NX_BATCH_CELL_LOOP(prefix, S, total)
    uout[pos] = subgrad_cell(uin + d.off, p + d.off, mask + d.off, x, y, d.W, d.H);
    mask_next[pos] = mask[pos] & MASK_PERSIST;
  }
}

__global__ void __launch_bounds__(256)
    kb_advect_dye(const SceneDev* __restrict__ sc, int S,
                  const int* __restrict__ prefix, int total,
                  const float* __restrict__ qin, float* __restrict__ qout,
                  const float2* __restrict__ u,
                  const uint8_t* __restrict__ mask, float dt) {
  NX_BATCH_CELL_LOOP(prefix, S, total)
    qout[pos] = advect_dye_cell(qin + d.off, u + d.off, mask + d.off, x, y,
                                d.W, d.H, d.W, d.H, dt, d.dye_speed, d.dye_diss);
  }
}

#undef NX_BATCH_CELL_LOOP

__global__ void __launch_bounds__(256)
    kb_fused(const SceneDev* __restrict__ sc, int S,
             const int* __restrict__ prefix, int total,
             const float2* __restrict__ uin, const uint8_t* __restrict__ mask,
             float2* __restrict__ uout, float* __restrict__ div,
             float* __restrict__ p1, float dt) {
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int s = find_scene(prefix, S, vb);
    const SceneDev d = sc[s];
    const int lb = vb - prefix[s];
    const int nbx = (d.W + fused::TX - 1) / fused::TX;
    fused::fused_vort_div_tile((lb % nbx) * fused::TX, (lb / nbx) * fused::TY,
                               uin + d.off, mask + d.off, uout + d.off,
                               div + d.off, p1 + d.off, d.W, d.H, dt, d.vort,
                               d.viscous, d.alpha, d.rbeta);
    __syncthreads();  // LDS is reused by the next tile
  }
}

template <int ROWS, int WAVES, int COLS>
__global__ void __launch_bounds__(WAVES * 64)
    kb_jacobi(const SceneDev* __restrict__ sc, int S,
              const int* __restrict__ prefix, int total,
              const float* __restrict__ pin, const float* __restrict__ b,
              const uint8_t* __restrict__ mask, float* __restrict__ pout,
              int k) {
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int s = find_scene(prefix, S, vb);
    const SceneDev d = sc[s];
    const int lb = vb - prefix[s];
    const int VX = 64 * COLS - 2 * k;
    const int nbx = (d.W + VX - 1) / VX;
    jt::jacobi_tb_tile<ROWS, WAVES, COLS>(lb % nbx, lb / nbx, pin + d.off,
                                          b + d.off, mask + d.off,
                                          pout + d.off, d.W, d.H, k, nullptr,
                                          0);
    __syncthreads();  // LDS is reused by the next tile
  }
}

template <typename K>
int resident_grid(K kernel, int block_threads, int total) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess || cus <= 0)
      cus = 256;
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block_threads, 0) !=
          hipSuccess || occ <= 0)
    occ = 1;
  const long long cap = (long long)cus * occ;
  return (int)std::max(1LL, std::min<long long>(cap, total));
}

}  // namespace

// ---- launchers used by the batch object --------------------------------
hipError_t launch_batch_jacobi(const SceneDev* sc, int S, const int* prefix,
                               int total, const float* pin, const float* b,
                               const uint8_t* mask, float* pout, JacobiCfg c,
                               hipStream_t s) {
#define NX_BJ(R, WV, CL)                                                       \
  if (c.rows == R && c.waves == WV && c.cols == CL) {                          \
    const int g = resident_grid(kb_jacobi<R, WV, CL>, WV * 64, total);         \
    kb_jacobi<R, WV, CL><<<g, WV * 64, 0, s>>>(sc, S, prefix, total, pin, b,   \
                                               mask, pout, c.k);               \
    return hipGetLastError();                                                  \
  }
  NX_BJ(4, 8, 1)
  NX_BJ(8, 8, 1)
  NX_BJ(8, 16, 1)
  NX_BJ(16, 8, 1)
  NX_BJ(2, 8, 2)
  NX_BJ(2, 16, 2)
  NX_BJ(4, 4, 2)
  NX_BJ(4, 8, 2)
  NX_BJ(4, 16, 2)
  NX_BJ(8, 4, 2)
  NX_BJ(8, 8, 2)
#undef NX_BJ
  return hipErrorInvalidValue;
}

}  // namespace nx

using namespace nx;

// ---------------------------------------------------------------- object
static int bfail(int code, const std::string& m) { return nx::set_error(code, m); }

#define NXB_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return bfail(NX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define NXB_TRY(expr)             \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != NX_OK) return rc_; \
  } while (0)

struct nx_batch {
  int S = 0;
  std::vector<int> W, H;
  std::vector<long long> off;
  long long cells = 0;
  hipStream_t stream = nullptr;
  float2* vel[2] = {nullptr, nullptr};
  float* p[2] = {nullptr, nullptr};
  float* div = nullptr;
  uint8_t* mask[2] = {nullptr, nullptr};
  float* dye[2] = {nullptr, nullptr};
  int vcur = 0, pcur = 0, mcur = 0, dcur = 0;
  std::vector<SceneDev> scenes;
  SceneDev* d_scenes = nullptr;
  bool scenes_dirty = true;
  std::vector<Splat> splats;  // sorted by scene
  Splat* d_splats = nullptr;
  size_t splat_cap = 0;
  // prefix tables: key = tile width * 100000 + tile height
  std::map<long long, std::pair<int*, int>> prefix;
  nx_params prm;
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timed;
  bool graph_on = false;
  hipGraphExec_t graph[8] = {};
  float graph_dt[8] = {};
  nx_params graph_prm[8];
};

extern "C" int nx_batch_destroy(nx_batch* b);

static int get_prefix(nx_batch* b, int tw, int th, int** dptr, int* total) {
  const long long key = (long long)tw * 100000 + th;
  auto it = b->prefix.find(key);
  if (it != b->prefix.end()) {
    *dptr = it->second.first;
    *total = it->second.second;
    return NX_OK;
  }
  std::vector<int> pre(b->S + 1, 0);
  for (int s = 0; s < b->S; ++s) {
    const long long n = (long long)((b->W[s] + tw - 1) / tw) * ((b->H[s] + th - 1) / th);
    if (pre[s] + n > (1ll << 30)) return bfail(NX_ERR_ARG, "batch too large");
    pre[s + 1] = pre[s] + (int)n;
  }
  int* d = nullptr;
  NXB_HIP(hipMalloc(&d, pre.size() * sizeof(int)));
  NXB_HIP(hipMemcpy(d, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
  b->prefix[key] = {d, pre[b->S]};
  *dptr = d;
  *total = pre[b->S];
  return NX_OK;
}

static void fill_scene(SceneDev& d, const nx_params& P) {
  d.speed = P.speed;
  d.diss = P.dissipation;
  d.vort = P.vorticity;
  d.viscous = P.viscosity > 0.0f;
  d.alpha = 0.0f;
  d.rbeta = 0.0f;
  if (d.viscous) {
    const double a = 1.0 / (double)P.viscosity;  // fluid_simulator.py:328-329
    d.alpha = (float)a;
    d.rbeta = (float)(1.0 / (4.0 + a));
  }
  d.bc = BcParams{P.bc_mode, P.has_borders, P.inflow};
  d.dye_speed = P.dye_speed;
  d.dye_diss = P.dye_dissipation;
}

static int sync_scenes(nx_batch* b) {
  if (!b->scenes_dirty) return NX_OK;
  NXB_HIP(hipStreamSynchronize(b->stream));
  NXB_HIP(hipMemcpy(b->d_scenes, b->scenes.data(), b->S * sizeof(SceneDev),
                    hipMemcpyHostToDevice));
  b->scenes_dirty = false;
  return NX_OK;
}

struct BTime {
  nx_batch* b;
  int tag;
  hipEvent_t a = nullptr, e = nullptr;
  BTime(nx_batch* b_, int t) : b(b_), tag(t) {
    if (!b->timing) return;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&e);
    (void)hipEventRecord(a, b->stream);
  }
  void end() {
    if (!b->timing) return;
    (void)hipEventRecord(e, b->stream);
    b->timed.push_back({tag, {a, e}});
  }
};

#define NXB_TIMED(b, tag, launch) \
  do {                            \
    BTime t_(b, tag);             \
    NXB_HIP(launch);              \
    t_.end();                     \
  } while (0)

#define NXB_KERNEL(b, tag, kernel, total, ...)                        \
  do {                                                                \
    const int g_ = resident_grid(kernel, 256, total);                 \
    BTime t_(b, tag);                                                 \
    kernel<<<g_, dim3(64, 4), 0, (b)->stream>>>(__VA_ARGS__);         \
    NXB_HIP(hipGetLastError());                                       \
    t_.end();                                                         \
  } while (0)

static int batch_jacobi(nx_batch* b, int sweeps, JacobiCfg c) {
  while (sweeps > 0) {
    JacobiCfg cc = c;
    if (sweeps < cc.k) cc.k = sweeps;
    int* pre;
    int total;
    NXB_TRY(get_prefix(b, 64 * cc.cols - 2 * cc.k, cc.rows * cc.waves - 2 * cc.k,
                       &pre, &total));
    NXB_TIMED(b, NX_TIME_JACOBI,
              launch_batch_jacobi(b->d_scenes, b->S, pre, total, b->p[b->pcur],
                                  b->div, b->mask[b->mcur], b->p[1 - b->pcur], cc,
                                  b->stream));
    b->pcur ^= 1;
    sweeps -= cc.k;
  }
  return NX_OK;
}

static int batch_step(nx_batch* b, float dt) {
  const nx_params& P = b->prm;
  int *pc, *pf;
  int tc, tf;
  NXB_TRY(get_prefix(b, 64, 4, &pc, &tc));
  NXB_TRY(get_prefix(b, fused::TX, fused::TY, &pf, &tf));
  const int a = b->vcur;
  float* dq = b->dye[0] ? b->dye[b->dcur] : nullptr;
  if (!b->splats.empty())
    NXB_KERNEL(b, NX_TIME_SPLAT, kb_splats, tc, b->d_scenes, b->S, pc, tc,
               b->vel[a], dq, b->d_splats);
  NXB_KERNEL(b, NX_TIME_ADVECT, kb_advect, tc, b->d_scenes, b->S, pc, tc,
             b->vel[a], b->vel[1 - a], b->mask[b->mcur], dt);
  NXB_KERNEL(b, NX_TIME_VORT_DIV, kb_fused, tf, b->d_scenes, b->S, pf, tf,
             b->vel[1 - a], b->mask[b->mcur], b->vel[a], b->div, b->p[0], dt);
  b->pcur = 0;
  JacobiCfg c{P.jacobi_k, P.jacobi_rows, P.jacobi_waves, P.jacobi_cols};
  NXB_TRY(batch_jacobi(b, P.iterations - 1, c));
  NXB_KERNEL(b, NX_TIME_PROJECT, kb_subgrad, tc, b->d_scenes, b->S, pc, tc,
             b->vel[a], b->p[b->pcur], b->mask[b->mcur], b->vel[1 - a],
             b->mask[1 - b->mcur]);
  b->vcur = 1 - a;
  b->mcur ^= 1;
  if (b->dye[0]) {
    NXB_KERNEL(b, NX_TIME_DYE, kb_advect_dye, tc, b->d_scenes, b->S, pc, tc,
               b->dye[b->dcur], b->dye[1 - b->dcur], b->vel[b->vcur],
               b->mask[b->mcur], dt);
    b->dcur ^= 1;
  }
  return NX_OK;
}

static int check_scene(nx_batch* b, int scene) {
  if (!b) return bfail(NX_ERR_ARG, "null batch");
  if (scene < 0 || scene >= b->S) return bfail(NX_ERR_ARG, "scene index out of range");
  return NX_OK;
}

extern "C" {

int nx_batch_create(int n, const int* widths, const int* heights, int with_dye,
                    nx_batch** out) {
  if (!out || n < 1 || !widths || !heights) return bfail(NX_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  nx_batch* b = new nx_batch();
  b->S = n;
  long long cells = 0;
  for (int s = 0; s < n; ++s) {
    if (widths[s] < 3 || heights[s] < 3 || (long long)widths[s] * heights[s] > (1ll << 28)) {
      delete b;
      return bfail(NX_ERR_ARG, "scene size out of range (each side >= 3)");
    }
    b->W.push_back(widths[s]);
    b->H.push_back(heights[s]);
    b->off.push_back(cells);
    cells += (long long)widths[s] * heights[s];
  }
  b->cells = cells;
  nx_default_params(&b->prm);
  b->scenes.resize(n);
  for (int s = 0; s < n; ++s) {
    SceneDev& d = b->scenes[s];
    memset(&d, 0, sizeof(d));
    d.W = b->W[s];
    d.H = b->H[s];
    d.off = b->off[s];
    fill_scene(d, b->prm);
  }
  auto cleanup = [&](int rc) {
    nx_batch_destroy(b);
    return rc;
  };
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  for (int i = 0; i < 2 && ok; ++i) {
    ok &= hipMalloc(&b->vel[i], cells * sizeof(float2)) == hipSuccess;
    ok &= hipMalloc(&b->p[i], cells * sizeof(float)) == hipSuccess;
    ok &= hipMalloc(&b->mask[i], cells) == hipSuccess;
    if (with_dye) ok &= hipMalloc(&b->dye[i], cells * sizeof(float)) == hipSuccess;
  }
  ok = ok && hipMalloc(&b->div, cells * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc(&b->d_scenes, n * sizeof(SceneDev)) == hipSuccess;
  if (!ok) return cleanup(bfail(NX_ERR_HIP, "batch allocation failed"));
  for (int i = 0; i < 2; ++i) {
    ok &= hipMemsetAsync(b->vel[i], 0, cells * sizeof(float2), b->stream) == hipSuccess;
    ok &= hipMemsetAsync(b->p[i], 0, cells * sizeof(float), b->stream) == hipSuccess;
    ok &= hipMemsetAsync(b->mask[i], 0, cells, b->stream) == hipSuccess;
    if (with_dye) ok &= hipMemsetAsync(b->dye[i], 0, cells * sizeof(float), b->stream) == hipSuccess;
  }
  ok &= hipMemsetAsync(b->div, 0, cells * sizeof(float), b->stream) == hipSuccess;
  ok &= hipStreamSynchronize(b->stream) == hipSuccess;
  if (!ok) return cleanup(bfail(NX_ERR_HIP, "batch initialisation failed"));
  *out = b;
  return NX_OK;
}

int nx_batch_destroy(nx_batch* b) {
  if (!b) return NX_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& g : b->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& t : b->timed) {
    (void)hipEventDestroy(t.second.first);
    (void)hipEventDestroy(t.second.second);
  }
  for (int i = 0; i < 2; ++i) {
    (void)hipFree(b->vel[i]);
    (void)hipFree(b->p[i]);
    (void)hipFree(b->mask[i]);
    (void)hipFree(b->dye[i]);
  }
  (void)hipFree(b->div);
  (void)hipFree(b->d_scenes);
  (void)hipFree(b->d_splats);
  for (auto& kv : b->prefix) (void)hipFree(kv.second.first);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return NX_OK;
}

int nx_batch_set_params(nx_batch* b, const nx_params* P) {
  if (!b || !P) return bfail(NX_ERR_ARG, "null argument");
  if (nx_check_params(P) != NX_OK) return NX_ERR_ARG;
  if (P->tol > 0) return bfail(NX_ERR_ARG, "the batched step runs a fixed sweep count");
  if (P->jacobi_seg != 0)
    return bfail(NX_ERR_ARG, "the batched step runs the Jacobi tile kernel (jacobi_seg 0)");
  b->prm = *P;
  for (auto& d : b->scenes) fill_scene(d, *P);
  b->scenes_dirty = true;
  return NX_OK;
}

int nx_batch_set_scene_params(nx_batch* b, int scene, const nx_params* P) {
  NXB_TRY(check_scene(b, scene));
  if (!P) return bfail(NX_ERR_ARG, "null params");
  if (nx_check_params(P) != NX_OK) return NX_ERR_ARG;
  if (P->jacobi_seg != 0)
    return bfail(NX_ERR_ARG, "the batched step runs the Jacobi tile kernel (jacobi_seg 0)");
  fill_scene(b->scenes[scene], *P);
  b->scenes_dirty = true;
  return NX_OK;
}

int nx_batch_set_obstacle_map(nx_batch* b, int scene, const uint8_t* map) {
  NXB_TRY(check_scene(b, scene));
  if (!map) return bfail(NX_ERR_ARG, "null map");
  const size_t n = (size_t)b->W[scene] * b->H[scene];
  std::vector<uint8_t> m(n);
  uint8_t* dst = b->mask[b->mcur] + b->off[scene];
  NXB_HIP(hipStreamSynchronize(b->stream));
  NXB_HIP(hipMemcpy(m.data(), dst, n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i)
    m[i] = (uint8_t)((m[i] & ~MASK_PERSIST) | (map[i] ? MASK_PERSIST : 0));
  NXB_HIP(hipMemcpy(dst, m.data(), n, hipMemcpyHostToDevice));
  return NX_OK;
}

int nx_batch_add_circle_obstacle(nx_batch* b, int scene, float px, float py, float radius,
                                 int is_static) {
  NXB_TRY(check_scene(b, scene));
  (void)is_static;  // both branches of AddCircleObstacle.comp:28-35 agree
  NXB_HIP(launch_circle(b->mask[b->mcur] + b->off[scene], b->W[scene], b->H[scene], px, py,
                        radius, b->stream));
  return NX_OK;
}

int nx_batch_add_triangle_obstacle(nx_batch* b, int scene, float x1, float y1, float x2,
                                   float y2, float x3, float y3, int is_static) {
  NXB_TRY(check_scene(b, scene));
  NXB_HIP(launch_triangle(b->mask[b->mcur] + b->off[scene], b->W[scene], b->H[scene], x1, y1,
                          x2, y2, x3, y3, is_static ? 1 : 0, b->stream));
  return NX_OK;
}

int nx_batch_set_step_splats(nx_batch* b, const float* rec, int n) {
  if (!b || n < 0 || (n > 0 && !rec)) return bfail(NX_ERR_ARG, "bad splat list");
  std::vector<std::pair<int, Splat>> v;
  for (int i = 0; i < n; ++i) {
    const float* f = rec + 7 * i;
    const int s = (int)f[0];
    if (s < 0 || s >= b->S || (float)s != f[0]) return bfail(NX_ERR_ARG, "bad splat scene");
    if (!(f[5] > 0)) return bfail(NX_ERR_ARG, "splat radius must be > 0");
    v.push_back({s, Splat{f[1], f[2], f[3], f[4], f[5], f[6]}});
  }
  std::stable_sort(v.begin(), v.end(),
                   [](const std::pair<int, Splat>& x, const std::pair<int, Splat>& y) {
                     return x.first < y.first;
                   });
  b->splats.clear();
  for (auto& d : b->scenes) d.splat_begin = d.splat_end = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    SceneDev& d = b->scenes[v[i].first];
    if (d.splat_begin == d.splat_end) d.splat_begin = (int)i;
    d.splat_end = (int)i + 1;
    b->splats.push_back(v[i].second);
  }
  NXB_HIP(hipStreamSynchronize(b->stream));
  if (b->splats.size() > b->splat_cap) {
    (void)hipFree(b->d_splats);
    b->d_splats = nullptr;
    NXB_HIP(hipMalloc(&b->d_splats, b->splats.size() * sizeof(Splat)));
    b->splat_cap = b->splats.size();
  }
  if (!b->splats.empty())
    NXB_HIP(hipMemcpy(b->d_splats, b->splats.data(), b->splats.size() * sizeof(Splat),
                      hipMemcpyHostToDevice));
  b->scenes_dirty = true;
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  return NX_OK;
}

int nx_batch_update(nx_batch* b, float dt) {
  if (!b) return bfail(NX_ERR_ARG, "null batch");
  NXB_TRY(sync_scenes(b));
  const nx_params& P = b->prm;
  if (!b->graph_on || b->timing) return batch_step(b, dt);
  const bool dye = b->dye[0] != nullptr;
  const int par = b->vcur | (b->mcur << 1) | (dye ? (b->dcur << 2) : 0);
  if (!b->graph[par] || b->graph_dt[par] != dt ||
      memcmp(&b->graph_prm[par], &P, sizeof(P)) != 0) {
    if (b->graph[par]) NXB_HIP(hipGraphExecDestroy(b->graph[par]));
    b->graph[par] = nullptr;
    // make every prefix table exist before capture (no allocation inside)
    int* tmp;
    int tot;
    NXB_TRY(get_prefix(b, 64, 4, &tmp, &tot));
    NXB_TRY(get_prefix(b, fused::TX, fused::TY, &tmp, &tot));
    JacobiCfg c{P.jacobi_k, P.jacobi_rows, P.jacobi_waves, P.jacobi_cols};
    NXB_TRY(get_prefix(b, 64 * c.cols - 2 * c.k, c.rows * c.waves - 2 * c.k, &tmp, &tot));
    const int rem = (P.iterations - 1) % c.k;
    if (rem) NXB_TRY(get_prefix(b, 64 * c.cols - 2 * rem, c.rows * c.waves - 2 * rem, &tmp, &tot));
    hipGraph_t g;
    NXB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int vc = b->vcur, mc = b->mcur, dc = b->dcur, pc = b->pcur;
    int rc = batch_step(b, dt);
    hipError_t ec = hipStreamEndCapture(b->stream, &g);
    b->vcur = vc;
    b->mcur = mc;
    b->dcur = dc;
    b->pcur = pc;
    if (rc != NX_OK) return rc;
    if (ec != hipSuccess) return bfail(NX_ERR_HIP, "hipStreamEndCapture failed");
    hipError_t ei = hipGraphInstantiate(&b->graph[par], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      b->graph[par] = nullptr;
      return bfail(NX_ERR_HIP, "hipGraphInstantiate failed");
    }
    b->graph_dt[par] = dt;
    b->graph_prm[par] = P;
  }
  NXB_HIP(hipGraphLaunch(b->graph[par], b->stream));
  int pcn = 0, rem = P.iterations - 1;
  while (rem > 0) {
    rem -= std::min(rem, P.jacobi_k);
    pcn ^= 1;
  }
  b->pcur = pcn;
  b->vcur ^= 1;
  b->mcur ^= 1;
  if (dye) b->dcur ^= 1;
  return NX_OK;
}

int nx_batch_set_graph(nx_batch* b, int on) {
  if (!b) return bfail(NX_ERR_ARG, "null batch");
  NXB_HIP(hipStreamSynchronize(b->stream));
  b->graph_on = on != 0;
  return NX_OK;
}

int nx_batch_jacobi_blocked(nx_batch* b, int sweeps, int k, int rows, int waves,
                            int cols) {
  if (!b || sweeps < 0) return bfail(NX_ERR_ARG, "bad arguments");
  JacobiCfg c{k, rows, waves, cols};
  if (!jacobi_cfg_supported(c)) return bfail(NX_ERR_ARG, "unsupported Jacobi tile");
  NXB_TRY(sync_scenes(b));
  return batch_jacobi(b, sweeps, c);
}

static int batch_field(nx_batch* b, int scene, int field, char** ptr, size_t* bytes) {
  NXB_TRY(check_scene(b, scene));
  const size_t n = (size_t)b->W[scene] * b->H[scene];
  const long long o = b->off[scene];
  switch (field) {
    case NX_FIELD_VELOCITY:
      *ptr = (char*)(b->vel[b->vcur] + o);
      *bytes = n * sizeof(float2);
      return NX_OK;
    case NX_FIELD_PRESSURE:
      *ptr = (char*)(b->p[b->pcur] + o);
      *bytes = n * sizeof(float);
      return NX_OK;
    case NX_FIELD_DIVERGENCE:
      *ptr = (char*)(b->div + o);
      *bytes = n * sizeof(float);
      return NX_OK;
    case NX_FIELD_MASK:
      *ptr = (char*)(b->mask[b->mcur] + o);
      *bytes = n;
      return NX_OK;
    case NX_FIELD_DYE:
      if (!b->dye[0]) return bfail(NX_ERR_STATE, "dye is not enabled");
      *ptr = (char*)(b->dye[b->dcur] + o);
      *bytes = n * sizeof(float);
      return NX_OK;
  }
  return bfail(NX_ERR_ARG, "unknown field id");
}

int nx_batch_upload(nx_batch* b, int scene, int field, const void* src, size_t bytes) {
  char* ptr;
  size_t nb;
  NXB_TRY(batch_field(b, scene, field, &ptr, &nb));
  if (!src || bytes != nb) return bfail(NX_ERR_ARG, "byte count does not match the field");
  NXB_HIP(hipStreamSynchronize(b->stream));
  NXB_HIP(hipMemcpyAsync(ptr, src, nb, hipMemcpyHostToDevice, b->stream));
  NXB_HIP(hipStreamSynchronize(b->stream));
  return NX_OK;
}

int nx_batch_download(nx_batch* b, int scene, int field, void* dst, size_t bytes) {
  char* ptr;
  size_t nb;
  NXB_TRY(batch_field(b, scene, field, &ptr, &nb));
  if (!dst || bytes != nb) return bfail(NX_ERR_ARG, "byte count does not match the field");
  NXB_HIP(hipStreamSynchronize(b->stream));
  NXB_HIP(hipMemcpyAsync(dst, ptr, nb, hipMemcpyDeviceToHost, b->stream));
  NXB_HIP(hipStreamSynchronize(b->stream));
  return NX_OK;
}

int nx_batch_residuals(nx_batch* b, double* out) {
  if (!b || !out) return bfail(NX_ERR_ARG, "null argument");
  // per scene: FP64 sums of the exact residual of the current pressure,
  // reduced in a fixed order; out[6*s + {sum r, sum r^2, sum b, sum b^2,
  // fluid cells, relative residual}]
  int maxblk = 0;
  for (int s = 0; s < b->S; ++s) maxblk = std::max(maxblk, residual_blocks(b->W[s], b->H[s]));
  double* red = nullptr;
  double* sums = nullptr;
  NXB_HIP(hipMalloc(&red, (size_t)maxblk * 5 * sizeof(double) * b->S));
  NXB_HIP(hipMalloc(&sums, (size_t)8 * b->S * sizeof(double)));
  int rc = NX_OK;
  for (int s = 0; s < b->S && rc == NX_OK; ++s) {
    const long long o = b->off[s];
    double* rs = red + (size_t)maxblk * 5 * s;
    if (launch_sums(b->p[b->pcur] + o, b->div + o, b->mask[b->mcur] + o, b->W[s],
                    b->H[s], rs, b->stream) != hipSuccess ||
        launch_reduce(rs, residual_blocks(b->W[s], b->H[s]), 5, 0, sums + 8 * s, -1.0,
                      nullptr, 0, 0, b->stream) != hipSuccess)
      rc = bfail(NX_ERR_HIP, "batch residual launch failed");
  }
  std::vector<double> h(8 * b->S);
  if (rc == NX_OK &&
      (hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                      b->stream) != hipSuccess ||
       hipStreamSynchronize(b->stream) != hipSuccess))
    rc = bfail(NX_ERR_HIP, "batch residual copy failed");
  (void)hipFree(red);
  (void)hipFree(sums);
  if (rc != NX_OK) return rc;
  for (int s = 0; s < b->S; ++s)
    for (int j = 0; j < 6; ++j) out[6 * s + j] = h[8 * s + j];
  return NX_OK;
}

int nx_batch_synchronize(nx_batch* b) {
  if (!b) return bfail(NX_ERR_ARG, "null batch");
  NXB_HIP(hipStreamSynchronize(b->stream));
  return NX_OK;
}

int nx_batch_cells(nx_batch* b, long long* cells) {
  if (!b || !cells) return bfail(NX_ERR_ARG, "null argument");
  *cells = b->cells;
  return NX_OK;
}

int nx_batch_set_timing(nx_batch* b, int on) {
  if (!b) return bfail(NX_ERR_ARG, "null batch");
  NXB_HIP(hipStreamSynchronize(b->stream));
  b->timing = on != 0;
  return NX_OK;
}

int nx_batch_collect_timing(nx_batch* b, double* ms, int* launches) {
  if (!b || !ms || !launches) return bfail(NX_ERR_ARG, "null argument");
  NXB_HIP(hipStreamSynchronize(b->stream));
  for (int i = 0; i < NX_TIME_COUNT; ++i) {
    ms[i] = 0.0;
    launches[i] = 0;
  }
  for (auto& t : b->timed) {
    float e = 0.0f;
    NXB_HIP(hipEventElapsedTime(&e, t.second.first, t.second.second));
    if (t.first >= 0 && t.first < NX_TIME_COUNT) {
      ms[t.first] += e;
      launches[t.first] += 1;
    }
    (void)hipEventDestroy(t.second.first);
    (void)hipEventDestroy(t.second.second);
  }
  b->timed.clear();
  return NX_OK;
}

}  // extern "C"
