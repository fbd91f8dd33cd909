/* natrix.h -- C-ABI of the MI355X (gfx950) Natrix fluid-step library.
 *
 * The reference (fbertola/Natrix) drives its fluid step from Python through
 * bgfx: it sets uniforms, binds buffers to slots and dispatches one compute
 * shader per stage.  This library replaces that boundary: one opaque
 * simulator object per grid owns its device buffers and HIP stream, and
 * every reference operation is one exported function.  All functions return
 * NX_OK (0) or a negative status; nx_last_error() gives the message of the
 * calling thread's last failure.  Calls enqueue on the simulator's stream
 * and return without waiting, except nx_download, nx_residual,
 * nx_synchronize, nx_get_stats and the tolerance-driven nx_update.
 * A simulator is not thread-safe (the reference drives bgfx from one thread).
 */
#ifndef NATRIX_H
#define NATRIX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NX_OK 0
#define NX_ERR_ARG (-1)   /* invalid argument (the reference raises ValueError) */
#define NX_ERR_HIP (-2)   /* HIP runtime failure */
#define NX_ERR_STATE (-3) /* call not valid in the simulator's current state */

/* Field identifiers for nx_upload / nx_download / nx_field_ptr.  Layouts are
 * row-major, index y*W + x (shader.AddVelocity.comp:26). */
#define NX_FIELD_VELOCITY 0   /* float2 (x, y) per cell: the current velocity */
#define NX_FIELD_PRESSURE 1   /* float: the current pressure iterate */
#define NX_FIELD_DIVERGENCE 2 /* float: the Poisson right-hand side */
#define NX_FIELD_CURL 3       /* float: vorticity scratch */
#define NX_FIELD_MASK 4       /* uint8: obstacle bits, bit0 = float2 channel x,
                                 bit1 = channel y, bit2 = persistent map */
#define NX_FIELD_DYE 5        /* float on the dye grid */
#define NX_FIELD_PRESSURE32 6 /* float32 pressure the projection used
                                 (reduced storage: p + correction) */

/* Storage formats of velocity, pressure, divergence and dye (arithmetic is
 * float32 throughout; stores round to nearest even). */
#define NX_STORAGE_F32 0
#define NX_STORAGE_BF16 1
#define NX_STORAGE_F16 2

#define NX_BC_REFERENCE 0 /* zero ring when has_borders (InitBoundaries.comp) */
#define NX_BC_FREE_SLIP 1 /* normal velocity 0, tangential copied inward */
#define NX_BC_INFLOW 2    /* left inflow, right outflow, free-slip top/bottom */

typedef struct nx_sim nx_sim;

/* Parameters.  The first six mirror the reference's properties and class
 * flags (natrix/core/fluid_simulator.py:28-35, validation :58-111). */
typedef struct nx_params {
  float speed;           /* > 0, default 500 */
  float dissipation;     /* > 0, default 1 */
  float vorticity;       /* >= 0, default 0 */
  float viscosity;       /* >= 0, default 0.1; the viscosity sweep runs if > 0 */
  int iterations;        /* > 0, default 50: Jacobi sweeps per step */
  int has_borders;       /* default 1 */
  int bc_mode;           /* NX_BC_*, default NX_BC_REFERENCE */
  float inflow;          /* inflow speed for NX_BC_INFLOW, default 1 */
  float dye_speed;       /* dye back-trace speed (SmoothParticlesArea._speed) */
  float dye_dissipation; /* dye dissipation, default 1 */
  int dye_in_step;       /* 1 (default): nx_update advects the dye after the
                            step; 0: the caller runs nx_stage_advect_dye */
  int fused;             /* 1 (default): fused kernels; 0: one kernel per
                            reference shader in the reference order */
  int jacobi_k;          /* sweeps per Jacobi launch (temporal blocking) */
  int jacobi_rows;       /* register rows per wavefront of the Jacobi tile */
  int jacobi_waves;      /* wavefronts per Jacobi block */
  int jacobi_cols;       /* grid columns per lane of the Jacobi tile */
  double tol;            /* > 0: solve each step until the relative residual
                            is <= tol (then `iterations` is ignored) */
  int max_iters;         /* sweep cap of the tolerance-driven solve */
  int check_every;       /* sweeps between residual checks */
  int correction_sweeps; /* reduced storage: float32 residual-correction
                            sweeps after the solve (0..16, default 8) */
  int jacobi_seg;        /* 0 (default): the Jacobi tile kernel; > 0 (float32
                            single grids): the streaming kernel, each
                            workgroup walking jacobi_seg output rows of a
                            strip (shapes as nx_set_smoother_tile); the
                            same sweeps bitwise */
} nx_params;

/* Statistics of the last pressure solve. */
typedef struct nx_stats {
  int sweeps;          /* Jacobi sweeps applied to the final pressure */
  int converged;       /* tolerance-driven solve: 1 if rel residual <= tol */
  double residual;     /* last measured relative residual (FP64) */
  int residual_sweeps; /* iterate index that residual was measured on */
} nx_stats;

/* ---- lifetime (FluidSimulator.__init__ / destroy, fluid_simulator.py:37-48,
 *      476-515) ---- */
int nx_create(int width, int height, nx_sim** out);
/* as nx_create with a storage format (NX_STORAGE_*); reduced storage runs
 * the fused step only (no per-stage entry points, no tolerance solve) */
int nx_create_ex(int width, int height, int storage, nx_sim** out);
int nx_destroy(nx_sim* sim);
int nx_default_params(nx_params* out);
int nx_get_params(const nx_sim* sim, nx_params* out);
/* validates like the reference setters; returns NX_ERR_ARG on bad values */
int nx_set_params(nx_sim* sim, const nx_params* prm);
/* the same validation without a simulator (no GPU needed) */
int nx_check_params(const nx_params* prm);
/* allocate a dye field of (width x height) driven by this simulator's
 * velocity (SmoothParticlesArea, demo/smooth_particles_area.py:26-43) */
int nx_enable_dye(nx_sim* sim, int width, int height);

/* ---- forcing and obstacles ---- */
/* FluidSimulator.add_velocity (fluid_simulator.py:116-131,
 * shader.AddVelocity.comp): cone splat, then every cell clamped to [-1,1] */
int nx_add_velocity(nx_sim* sim, float px, float py, float vx, float vy,
                    float radius);
/* FluidSimulator.add_circle_obstacle (:135-153, AddCircleObstacle.comp) */
int nx_add_circle_obstacle(nx_sim* sim, float px, float py, float radius,
                           int is_static);
/* FluidSimulator.add_triangle_obstacle (:156-172, AddTriangleObstacle.comp) */
int nx_add_triangle_obstacle(nx_sim* sim, float x1, float y1, float x2,
                             float y2, float x3, float y3, int is_static);
/* SmoothParticlesArea.add_particles (smooth_particles_area.py:70-83,
 * shader.AddParticle.comp): cone dye splat clamped to [0,255] */
int nx_add_dye(nx_sim* sim, float px, float py, float amount, float radius);
/* the demo's dye display pass (demo/shaders/demo.ComputeShader.comp:9-19):
 * out[dye_h * dye_w] (device, one uint32 per cell, bytes R,G,B,A) = the dye
 * value stored to an rgba8 image in every channel: clamp(v, 0, 1) * 255
 * rounded to nearest (ties up), NaN -> 0; on the simulator's stream */
int nx_dye_rgba8(nx_sim* sim, void* out);
/* the demo's velocity arrows (demo/shaders/demo.QuiverFragmentShader.frag:
 * 14-70, uniforms as simulation_demo.py:251-273: WindowSize = the
 * framebuffer, VelocitySize = the simulator's grid, ArrowTileSize):
 * out[window_h * window_w] (device, one uint32 per pixel, bytes R,G,B,A =
 * 255, 255, 255, alpha) with row y at gl_FragCoord.y = y + 0.5; alpha =
 * 1 - clamp(distance to the pixel's tile arrow, 0, 1) rounded to 8 bits;
 * on the simulator's stream */
int nx_quiver_rgba8(nx_sim* sim, int window_w, int window_h, float arrow_tile, void* out);
/* Gaussian splats (no reference counterpart): n records of 6 floats
 * (px, py, ax, ay, radius, dye); velocity += (ax, ay) * w, dye += dye * w,
 * w = exp(-d^2 / r^2).  Amplitudes are final (already scaled by dt). */
int nx_add_gaussian_splats(nx_sim* sim, const float* splats, int n);
/* Persistent obstacle map (no reference counterpart): W*H bytes, non-zero =
 * solid.  Survives the end-of-step obstacle clear. */
int nx_set_obstacle_map(nx_sim* sim, const uint8_t* map);

/* ---- the step (FluidSimulator.update, fluid_simulator.py:174-280, then the
 *      dye advection SmoothParticlesArea.update, :85-104, if dye is on) ---- */
int nx_update(nx_sim* sim, float dt);
/* ---- the step in phases, for a caller-provided pressure guess (the
 *      learned solve).  begin: walls + advection + the fused curl ..
 *      divergence stage; the divergence is centred on its fluid mean and the
 *      pressure set to the first Jacobi sweep from zero.  The caller may then
 *      overwrite or correct the pressure (NX_FIELD_PRESSURE, or the coarse
 *      transfers below) and call nx_solve any number of times; finish:
 *      projection, obstacle clear, dye advection. ---- */
int nx_step_begin(nx_sim* sim, float dt);
/* Jacobi from the current pressure until the FP64 relative residual is
 * <= tol or max_sweeps sweeps were applied; `out` (nullable) gets the stats
 * of the whole step so far */
int nx_solve(nx_sim* sim, double tol, int max_sweeps, nx_stats* out);
/* coarse correction problem: coarse_rhs[Hc*Wc] (device, float32) = factor^2
 * x the fluid mean of the fine residual per factor x factor block,
 * coarse_mask[Hc*Wc] (device, uint8) = 1 for blocks with no fluid cell;
 * Wc = ceil(W/factor), Hc = ceil(H/factor) */
int nx_restrict_residual(nx_sim* sim, int factor, float* coarse_rhs,
                         uint8_t* coarse_mask);
/* the restriction of the face-weighted coarse operator (no reference
 * counterpart): coarse_rhs = the sum of the fine residual over the fluid
 * cells of each block (factor^2 x the block mean, solid cells counting as
 * 0), coarse_mask as nx_restrict_residual, faces[2*Hc*Wc] (device, float32)
 * = the open fraction of each coarse cell's east face (first Hc*Wc) and
 * south face: the fine rows (columns) of the block whose cells on both
 * sides of the block boundary are fluid, over factor; 0 at the grid's
 * edge and for solid cells */
int nx_restrict_residual_faces(nx_sim* sim, int factor, float* coarse_rhs,
                               uint8_t* coarse_mask, float* faces);
/* pressure += scale * bilinear interpolation of coarse[Hc*Wc] (device) on
 * fluid cells, weights renormalised over the fluid coarse cells of
 * coarse_mask (as written by nx_restrict_residual) */
int nx_prolong_add(nx_sim* sim, int factor, const float* coarse,
                   const uint8_t* coarse_mask, float scale);
/* pressure += a * d with d the interpolation of coarse[Hc*Wc] as in
 * nx_prolong_add and a = (d.r)/(d.Ld) the minimiser of the solve's energy
 * functional along d (r = b - L p, L the Jacobi operator; a = 0 if d has
 * no energy), so the error in the energy norm does not grow; all on the
 * simulator's stream.  out (host, may be NULL; then no synchronisation)
 * receives {a, d.r, d.Ld}. */
int nx_prolong_line_search(nx_sim* sim, int factor, const float* coarse,
                           const uint8_t* coarse_mask, double* out);
/* smoothing of the coarse correction problem L e = rc on a Wc x Hc grid
 * (at most 128 x 128; device buffers as written by nx_restrict_residual),
 * e[Hc*Wc] updated in place, one launch on the simulator's stream:
 * lambda_min <= 0: `iterations` Jacobi sweeps (the fine sweep's operator);
 * 0 < lambda_min < 8: `iterations` Chebyshev iterations of the same
 * operator on [lambda_min, 8] (-L has its spectrum in [0, 8]) */
int nx_coarse_smooth(nx_sim* sim, int Wc, int Hc, float* e, const float* rc,
                     const uint8_t* coarse_mask, int iterations, float lambda_min);
/* Chebyshev smoothing (0 < lambda_min < 8) of the face-weighted coarse
 * problem (no reference counterpart): (L e)_I = sum over the faces of I of
 * w (e_J - e_I) with the weights of nx_restrict_residual_faces (spectrum
 * of -L in [0, 8]); a neighbour across a face of weight w enters as the
 * centre (w = 0), itself (w = 1) or fma(w, e_J - e_I, e_I), summed in the
 * plain kernel's order */
int nx_coarse_smooth_faces(nx_sim* sim, int Wc, int Hc, float* e, const float* rc,
                           const uint8_t* coarse_mask, const float* faces, int iterations,
                           float lambda_min);
/* `sweeps` temporal-blocked Jacobi sweeps of the current pressure system,
 * then (omega > 0) one damped sweep, then the FP64 relative residual; one
 * synchronisation.  out (host, may be NULL) = {relative residual, the last
 * line search's step a}.  Counted in the step's sweeps. */
int nx_smooth(nx_sim* sim, int sweeps, float omega, double* out);
/* `iterations` Chebyshev iterations of the current pressure system on the
 * interval [lambda_min, 8] of -L (0 < lambda_min < 8; the three-term
 * recurrence of Saad, Iterative Methods for Sparse Linear Systems, Alg.
 * 12.1, from the current pressure): every mode of -L above lambda_min is
 * damped by 1 / T_n((8 + lambda_min) / (8 - lambda_min)) and none grows;
 * then the FP64 relative residual, as nx_smooth (same out).  tile (may be
 * NULL: the simulator's smoother tile, see below) = {stencil passes per
 * launch, rows per wave, waves per block, columns per lane} of the tile
 * kernel; shapes (2,16,2), (4,8,2), (4,16,2), (8,8,1), (8,16,1), at most
 * 32 passes, 64 cols - 2 passes >= 8 and rows waves - 2 passes >= 8.
 * Replaces the Jacobi smoothing of the learned cycle; no reference
 * counterpart (the reference smooths only by Jacobi,
 * shader.Poisson.comp:24-37).  Counted in the step's sweeps (one per
 * iteration). */
int nx_smooth_chebyshev(nx_sim* sim, int iterations, float lambda_min, const int* tile,
                        double* out);
/* the simulator's smoother tile, used when nx_smooth_chebyshev or
 * nx_learned_cycles get tile == NULL: n = 4 values as nx_smooth_chebyshev's
 * tile (tile kernel), or n = 5 with a fifth value seg > 0: the streaming
 * kernel, whose workgroups each walk `seg` output rows of a (64 cols)-column
 * strip in chunks of rows x waves rows (shapes (4,8,2), (2,16,2), (4,16,2),
 * (8,8,2); at most 16 passes per launch, 64 cols - 2 passes >= 8); the
 * same results bitwise.  Default {8, 4, 8, 2}.  nx_get_smoother_tile
 * writes the 5 values (seg 0 for the tile kernel).  No reference
 * counterpart (launch shape of the smoother). */
int nx_set_smoother_tile(nx_sim* sim, const int* tile, int n);
int nx_get_smoother_tile(nx_sim* sim, int* tile);
/* `count` damped Jacobi sweeps p <- p + omega (J(p) - p) of the current
 * pressure system (0 < omega <= 1; omega = 1/2 removes the checkerboard
 * modes the plain sweep cannot damp); counted in the step's sweeps */
int nx_damped_sweeps(nx_sim* sim, int count, float omega);
/* the same line search along an arbitrary direction d[H*W] (device,
 * float32): pressure += a * d on fluid cells */
int nx_line_search(nx_sim* sim, const float* d, double* out);
/* iterative refinement inside a step (for tolerances below the float32
 * round-off floor of a large pressure): the pressure becomes a float32 pair
 * ph + pl, initially 0.  nx_refine_residual evaluates r = b - A(ph + pl) in
 * FP64, returns its relative FP64 residual, and makes (r, p1 of A e = r)
 * the current right-hand side and pressure, so nx_solve / the coarse
 * transfers then work on the correction e; nx_refine_accumulate adds the
 * current pressure (e) to the pair with an error-free two-sum;
 * nx_refine_end (or nx_step_finish) sets the pressure to ph + pl and
 * restores the right-hand side. */
int nx_refine_begin(nx_sim* sim);
int nx_refine_residual(nx_sim* sim, double* rel);
int nx_refine_accumulate(nx_sim* sim);
int nx_refine_end(nx_sim* sim);
int nx_step_finish(nx_sim* sim);

/* Capture the fixed-iteration step into a HIP graph and replay it on later
 * nx_update calls with the same dt and parameters (on = 1). */
int nx_set_graph(nx_sim* sim, int on);

/* ---- per-stage entry points (one reference shader each), operating on the
 *      simulator's current buffers in the reference's slot semantics ---- */
int nx_stage_boundaries(nx_sim* sim);               /* InitBoundaries (+modes) */
int nx_stage_advect(nx_sim* sim, float dt);         /* AdvectVelocity + flip */
int nx_stage_curl(nx_sim* sim);                     /* CalcVorticity */
int nx_stage_confine(nx_sim* sim, float dt);        /* ApplyVorticity + flip */
int nx_stage_viscosity(nx_sim* sim);                /* Viscosity + flip */
int nx_stage_divergence(nx_sim* sim);               /* Divergence */
int nx_stage_clear_pressure(nx_sim* sim);           /* ClearBuffer on p */
int nx_stage_jacobi(nx_sim* sim, int sweeps);       /* Poisson + flip, xN */
int nx_stage_subtract_gradient(nx_sim* sim);        /* SubtractGradient + flip */
int nx_stage_clear_obstacles(nx_sim* sim);          /* ClearBuffer on obstacles */
int nx_stage_advect_dye(nx_sim* sim, float dt);     /* AdvectParticle + flip */
/* fused middle stage: curl, confinement, viscosity, divergence, p1 */
int nx_stage_fused_vort_div(nx_sim* sim, float dt);
/* `sweeps` Jacobi sweeps with the temporal-blocked kernel, k per launch;
 * seg = 0: the tile kernel, seg > 0: the streaming kernel (nx_params) */
int nx_jacobi_blocked(nx_sim* sim, int sweeps, int k, int rows, int waves,
                      int cols, int seg);

/* ---- state access ---- */
int nx_upload(nx_sim* sim, int field, const void* src, size_t bytes);
int nx_download(nx_sim* sim, int field, void* dst, size_t bytes);
/* device pointer of the current buffer of `field` (borrowed: valid until
 * the next call that flips that field) -- get_velocity_buffer,
 * fluid_simulator.py:113-114 */
int nx_field_ptr(nx_sim* sim, int field, void** dev_ptr, size_t* bytes);
/* FP64 relative residual of the current pressure against the current
 * divergence: out[6] = {sum r, sum r^2, sum b, sum b^2, fluid cells, rel} */
int nx_residual(nx_sim* sim, double* out);
int nx_get_stats(nx_sim* sim, nx_stats* out);
int nx_synchronize(nx_sim* sim);
/* the simulator's HIP stream (hipStream_t) */
int nx_get_stream(nx_sim* sim, void** stream);
/* grid size (FluidSimulator.width / height, fluid_simulator.py:37-48) */
int nx_get_size(const nx_sim* sim, int* width, int* height);

/* ---- timing: HIP events around every launch of the selected kernels ---- */
#define NX_TIME_ADVECT 0
#define NX_TIME_VORT_DIV 1
#define NX_TIME_JACOBI 2
#define NX_TIME_PROJECT 3
#define NX_TIME_DYE 4
#define NX_TIME_SPLAT 5
#define NX_TIME_CHEB 6 /* fine Chebyshev smoothing (nx_smooth_chebyshev) */
#define NX_TIME_COUNT 7
int nx_set_timing(nx_sim* sim, int on);
/* synchronises; per tag: total ms and number of launches since the last
 * reset (ms, launches: NX_TIME_COUNT entries each); resets the
 * accumulators */
int nx_collect_timing(nx_sim* sim, double* ms, int* launches);

/* ---- batched scenes: many independent grids of uneven sizes advanced by
 *      one persistent launch per stage (no reference counterpart; each
 *      scene behaves as its own nx_sim with reference-mode semantics) ---- */
typedef struct nx_batch nx_batch;
/* n scenes of widths[i] x heights[i] (each side >= 3); dye on the velocity
 * grid when with_dye */
int nx_batch_create(int n, const int* widths, const int* heights, int with_dye,
                    nx_batch** out);
int nx_batch_destroy(nx_batch* b);
/* solver settings (iterations, Jacobi tile) and every scene's parameters */
int nx_batch_set_params(nx_batch* b, const nx_params* prm);
/* one scene's parameter block (speed, dissipation, vorticity, viscosity,
 * walls, inflow, dye speed/dissipation) */
int nx_batch_set_scene_params(nx_batch* b, int scene, const nx_params* prm);
int nx_batch_set_obstacle_map(nx_batch* b, int scene, const uint8_t* map);
/* per-step obstacles of one scene, as nx_add_circle_obstacle /
 * nx_add_triangle_obstacle (fluid_simulator.py:135-172): marked until the
 * end-of-step clear of the next nx_batch_update */
int nx_batch_add_circle_obstacle(nx_batch* b, int scene, float px, float py, float radius,
                                 int is_static);
int nx_batch_add_triangle_obstacle(nx_batch* b, int scene, float x1, float y1, float x2,
                                   float y2, float x3, float y3, int is_static);
/* Gaussian splats applied at the start of every nx_batch_update: n records
 * of 7 floats (scene, px, py, ax, ay, radius, dye) */
int nx_batch_set_step_splats(nx_batch* b, const float* rec, int n);
int nx_batch_update(nx_batch* b, float dt);
int nx_batch_set_graph(nx_batch* b, int on);
int nx_batch_jacobi_blocked(nx_batch* b, int sweeps, int k, int rows, int waves,
                            int cols);
int nx_batch_upload(nx_batch* b, int scene, int field, const void* src,
                    size_t bytes);
int nx_batch_download(nx_batch* b, int scene, int field, void* dst,
                      size_t bytes);
int nx_batch_cells(nx_batch* b, long long* cells);
/* FP64 relative residual of every scene's current pressure (synchronises):
 * out[6*s + {sum r, sum r^2, sum b, sum b^2, fluid cells, rel}] */
int nx_batch_residuals(nx_batch* b, double* out);
int nx_batch_synchronize(nx_batch* b);
int nx_batch_set_timing(nx_batch* b, int on);
int nx_batch_collect_timing(nx_batch* b, double* ms, int* launches);

/* ---- U-Net building blocks (learned pressure guess; no reference
 *      counterpart).  NHWC bfloat16 activations, KRSC bfloat16 3x3 weights
 *      (w[cout][ky*3+kx][cin]), float32 accumulation on the matrix cores.
 *      All pointers are device pointers; `stream` is a hipStream_t. ---- */
/* y = [relu](conv3x3(x, w) + bias), zero padding; cin % 8 == 0, cout % 32 == 0 */
int nx_conv3x3_fwd(const void* x, const void* w, const float* bias, void* y, int n,
                   int h, int wd, int cin, int cout, int relu, void* stream);
/* the same convolution with the 9 cin / 16 K steps split over `splits`
 * groups, their float32 partial sums summed in split order with bias and
 * ReLU: more workgroups for small images with many channels (cin % 16 ==
 * 0).  When every split fits one wave of a workgroup (at most 16 splits of
 * at most 18 K steps) the splits of a 32 x 32 output tile run in one
 * workgroup and are summed in its LDS; otherwise they run as separate
 * workgroups that write work[splits * n*h*w * cout] (device, float32),
 * summed by a second launch.  Both give the same values. */
int nx_conv3x3_fwd_split(const void* x, const void* w, const float* bias, void* y,
                         float* work, int splits, int n, int h, int wd, int cin, int cout,
                         int relu, void* stream);
/* dx = conv3x3 input gradient of dy; wt: scratch of 9*cin*cout bfloat16 */
int nx_conv3x3_dgrad(const void* dy, const void* w, void* dx, void* wt, int n, int h,
                     int wd, int cin, int cout, void* stream);
/* dw (float32 [cout][9][cin]) = weight gradient; work: splits*cout*9*cin float32 */
int nx_conv3x3_wgrad(const void* x, const void* dy, float* dw, float* work, int splits,
                     int n, int h, int wd, int cin, int cout, void* stream);
/* dz = dy where y > 0 else 0 (bfloat16, `count` elements) */
int nx_relu_bwd(const void* y, const void* dy, void* dz, long long count, void* stream);
/* out[c] = sum over rows of a bfloat16 (rows x c) matrix; work: 1024*c float32 */
int nx_colsum(const void* a, float* out, float* work, long long rows, int c,
              void* stream);
/* 2x2 average pool (backward: spread dy/4); h, wd = the full-resolution size */
int nx_avgpool2(const void* x, void* y, int n, int h, int wd, int c, int backward,
                void* stream);
/* 2x nearest upsample (backward: sum of 4); h, wd = the full-resolution size */
int nx_upsample2(const void* x, void* y, int n, int h, int wd, int c, int backward,
                 void* stream);
/* channel concat y = [a, b] (split = 1: a, b = y's parts) */
int nx_concat(void* y, void* a, void* b, long long pixels, int ca, int cb, int split,
              void* stream);
/* 1x1 head to one float32 channel, and its backward (dwb[0..c-1] = dw,
 * dwb[c] = dbias; work: 1024*(c+1) float32) */
int nx_head_fwd(const void* x, const float* w, const float* bias, float* out,
                long long pixels, int c, void* stream);
int nx_head_bwd(const void* x, const float* w, const float* dout, void* dx, float* dwb,
                float* work, long long pixels, int c, void* stream);

/* ---- native learned correction (inference; no reference counterpart).
 *      The U-Net of natrix_amd.learned.unet on one coarse grid of H x W
 *      (multiples of 8, at most 128 x 128): conv_w[14] / conv_b[14] are the 3x3 convolutions in
 *      forward order (e0 a, e0 b, e1 a, e1 b, e2 a, e2 b, e3 a, e3 b, d2 a,
 *      d2 b, d1 a, d1 b, d0 a, d0 b) as bfloat16 KRSC weights and float32
 *      biases, head_w[base] / head_b[1] the 1x1 head, all device pointers
 *      the caller keeps alive (biases and head are read at every forward);
 *      the runner copies the convolution weights during nx_unet_create (on
 *      the null stream, which it synchronises; tiled for coalesced loads
 *      where cin % 16 == 0), so they must be complete when it is called,
 *      and owns its activation buffers. ---- */
typedef struct nx_unet nx_unet;
int nx_unet_create(int H, int W, int base, const void* const* conv_w,
                   const float* const* conv_b, const float* head_w, const float* head_b,
                   nx_unet** out);
int nx_unet_destroy(nx_unet* net);
/* out[H*W] (float32) = network(x[H*W*8] bfloat16 NHWC), batch 1 */
int nx_unet_forward(nx_unet* net, const void* x, float* out, void* stream);
/* one learned correction of the simulator's pressure, on its stream, no
 * synchronisation: restrict the residual (factor), normalise to unit RMS,
 * the network, rescale, `coarse_iters` Chebyshev iterations on
 * [lambda_min, 8] (Jacobi sweeps if lambda_min <= 0), remove the fluid
 * mean, interpolate and add with the energy line search (out as for
 * nx_prolong_line_search).  The coarse-grid part (normalisation to mean
 * removal) is captured as a HIP graph on the second call with the same
 * stream and smoothing parameters and replayed from then on. */
int nx_learned_correct(nx_sim* sim, nx_unet* net, int factor, int coarse_iters,
                       float lambda_min, double* out);
/* the learned cycles of one correction solve, with one host call: up to
 * max_cycles times nx_learned_correct (factor, coarse_iters, lambda_min)
 * followed by smooth_iters fine iterations (nx_smooth_chebyshev on
 * [fine_lambda_min, 8] with `tile`, or nx_smooth with smooth_iters - 1
 * Jacobi sweeps and one damped sweep when fine_lambda_min <= 0), stopping
 * when the FP64 relative residual is <= target (converged), when a line
 * search takes step 0 or a cycle leaves more than stall x the previous
 * cycle's residual (abandoned), or after max_cycles.  result[4] (host) =
 * {cycles run, cycles whose step moved the pressure, converged, abandoned};
 * records (host, may be NULL) holds records_capacity doubles, at least
 * 3 * max_cycles (else NX_ERR_ARG): per cycle {line-search step, the
 * step's sweeps so far, residual}.  All arguments, the smoothing tile
 * (cheb tiles as nx_smooth_chebyshev) and 0 < fine_lambda_min < 8 included,
 * are checked before anything is enqueued.  One synchronisation per
 * cycle.  No reference counterpart (the solver loop of the learned step). */
int nx_learned_cycles(nx_sim* sim, nx_unet* net, int factor, int coarse_iters,
                      float lambda_min, int smooth_iters, float fine_lambda_min,
                      const int* tile, double target, double stall, int max_cycles,
                      double* records, int records_capacity, int* result);
/* the runner's coarse buffers (rhs, mask, correction), for inspection */
int nx_unet_coarse(nx_unet* net, float** rc, uint8_t** mc, float** e);
/* bypass != 0: the runner's corrections start the coarse smoothing from
 * zero instead of the network's output (the same cycle without the
 * network, to measure what the network contributes); 0 restores it.  No
 * reference counterpart. */
int nx_unet_set_bypass(nx_unet* net, int bypass);
/* three-level cycle (no reference counterpart): with mid_factor >= 2 each
 * correction with Chebyshev coarse smoothing restricts to a middle grid
 * mid_factor times coarser than the fine one (block sums and face
 * weights), from there to the network's grid (the correction's factor must
 * be a multiple of mid_factor), and smooths the middle problem with
 * mid_iters Chebyshev iterations of its face-weighted operator on
 * [mid_lambda_min, 8] before interpolating to the fine grid with the line
 * search; mid_factor = 0 restores the two-grid cycle */
int nx_unet_set_levels(nx_unet* net, int mid_factor, int mid_iters, float mid_lambda_min);

/* ---- FP64 oracle solve on the GPU (conjugate gradients on the same
 *      operator, fluid cells, right-hand side centred on its fluid mean) ---- */
/* b: float32 device field, mask: uint8 device field, x: W*H doubles (device);
 * out[0] = true relative residual of x, out[1] = iterations */
int nx_fp64_solve(const float* b, const uint8_t* mask, int W, int H, double* x,
                  double tol, int max_iter, double* out, void* stream);
/* FP64 conjugate gradients on `count` independent n x n problems
 * (n * n <= 16384) stored back to back: b (float32) and mask (uint8, != 0
 * solid) of count * n * n cells each, device.  One workgroup per problem
 * runs the whole solve (same operator, centring and mean projection as
 * nx_fp64_solve) until the relative residual is <= tol or max_iter
 * iterations.  xf (float32, device) gets each problem's solution with its
 * fluid mean removed (0 on solid cells), residuals[count] and
 * iterations[count] (device) the true relative residual of the FP64
 * solution and the iterations run; work: count * n * n doubles (device).
 * Used to make the learned guess's training targets; no reference
 * counterpart. */
int nx_fp64_solve_batch(const float* b, const uint8_t* mask, int n, int count, double* work,
                        float* xf, double tol, int max_iter, double* residuals,
                        int* iterations, void* stream);
/* nx_fp64_solve with no allocation: the same kernels, in the same order, on
 * a static device array of the library (about 6 MiB, held on every GPU the
 * library runs on until the process ends) instead of hipMalloc'd buffers,
 * so the results are bitwise those of nx_fp64_solve.  Grids of at most
 * 512 x 512 = 262144 cells (NX_ERR_ARG beyond).  Calls on one GPU share
 * that array: calls that overlap on one GPU race; a call does not depend on
 * what a previous call left there.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under stream capture,
 * serving loops, harnesses that forbid it). */
int nx_fp64_solve_noalloc(const float* b, const uint8_t* mask, int W, int H, double* x,
                          double tol, int max_iter, double* out, void* stream);
/* relative L2 error of the mean-free float32 p against the mean-free x */
int nx_fp64_compare(const float* p, const double* x, const uint8_t* mask, int W, int H,
                    double* rel, void* stream);

/* attainable bandwidth: a plain device copy of `bytes` (multiple of 16),
 * `iters` times; *gbps = (read + write bytes) / time.  nx_copy_bandwidth
 * runs the counter-calibration copy (one float4 per thread per grid-stride
 * iteration); nx_copy_bandwidth_variant runs variant 0-8 (0 = that copy,
 * 1/2 = 4/8 float4 in flight per thread, 3 = 8 with non-temporal loads and
 * stores, 4 = one contiguous chunk per workgroup, 5/6/7/8 = 4 accesses of
 * 8/4/2/1 bytes per thread) or, with variant -1, all of them and reports
 * the best.  No reference counterpart (measurement). */
int nx_copy_bandwidth(size_t bytes, int iters, double* gbps);
int nx_copy_bandwidth_variant(size_t bytes, int iters, int variant, double* gbps);

const char* nx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NATRIX_H */
