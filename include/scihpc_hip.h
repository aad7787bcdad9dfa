/*
 * scihpc_hip.h -- C ABI of the MI355X (gfx950) projection-step hot path.
 *
 * Plain C: opaque handles, host or device pointers plus sizes, int status
 * codes.  No torch, HIP or C++ types cross this boundary.  Every entry point
 * names the SciHPC interface it replaces (paths relative to the SciHPC
 * source tree).  Handles are not thread safe; all work of one handle is
 * ordered on that handle's HIP stream, and calls that return host data or
 * host scalars synchronise that stream.
 *
 * Array layouts
 *   MG layout     : unpadded, x fastest, idx = i + nx*(j + ny*k), exactly
 *                   multigrid_base::of() (multigrid_base.cpp:8-14).
 *   padded layout : scalar_data's data[i][j][k] made contiguous, k fastest,
 *                   ghost width 3 on every axis < ndim
 *                   (scalar_data.cpp:73-158, global.cpp:7-19).
 */
#ifndef SCIHPC_HIP_H
#define SCIHPC_HIP_H

#if defined(__GNUC__) || defined(__clang__)
#define SCIHPC_API __attribute__((visibility("default")))
#else
#define SCIHPC_API
#endif

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference has none: assert only, bc_factory.cpp:8-21) */
enum {
  SCIHPC_OK = 0,
  SCIHPC_NOT_CONVERGED = 1,
  SCIHPC_BAD_ARG = 2,
  SCIHPC_HIP_ERROR = 3,
  SCIHPC_FELL_BACK = 4,
  SCIHPC_IO_ERROR = 5
};

/* multigrid boundary operator: init_NeumannBC / init_DirichletBC
 * (multigrid_base.cpp:236-345) */
enum { SCIHPC_MG_NEUMANN = 0, SCIHPC_MG_DIRICHLET = 1 };

/* precision of the inner V-cycle */
enum { SCIHPC_PREC_FP64 = 0, SCIHPC_PREC_MIXED = 1 };

/* boundary condition types, same meaning as bc_factory.h:18-22 */
enum {
  SCIHPC_BC_NEUMANN = 0,
  SCIHPC_BC_DIRICHLET = 1,
  SCIHPC_BC_PERIODIC = 2,
  SCIHPC_BC_NO_SLIP = 3,
  SCIHPC_BC_SLIP = 4
};

/* last error message of the calling thread ("" when none) */
SCIHPC_API const char *scihpc_last_error(void);

/* Zeroed host memory for fields that cross this ABI on every call:
 * page-locked (hipHostMalloc), so that their copies run at the link's rate
 * without staging, or ordinary memory when no page-locked memory is left.
 * Free with scihpc_host_free.  No counterpart in the reference (its
 * arrays are new[]'d, global.cpp:7-28); for the facade's field types. */
SCIHPC_API int scihpc_host_alloc(size_t bytes, void **out);
SCIHPC_API int scihpc_host_free(void *p);

/* ------------------------------------------------------------------ */
/* Pressure-Poisson multigrid                                          */
typedef struct scihpc_mg scihpc_mg;

/* multigrid::multigrid(scalar_data*, structured_grid*) (multigrid.cpp:81-122)
 * plus shared_solvers_mg_init_Neumann/_Dirichlet (shared_solvers.cpp:17-28):
 * builds the level list with the reference's coarsening rule (ratio 2, else
 * 3, while every coarse size stays > 3, multigrid.cpp:8-41) and the
 * constant-coefficient operator of every level.  ndim 2 ignores nz/dz. */
SCIHPC_API int scihpc_mg_create(int ndim, int nx, int ny, int nz, double dx, double dy,
                     double dz, int bc, scihpc_mg **out);
SCIHPC_API int scihpc_mg_destroy(scihpc_mg *h);
/* multigrid::level_num (multigrid.h:21) */
SCIHPC_API int scihpc_mg_num_levels(const scihpc_mg *h, int *levels);
/* multigrid_base::{nx,ny,nz,degree} of at[level] (multigrid_base.h:31) */
SCIHPC_API int scihpc_mg_level_shape(const scihpc_mg *h, int level, int *nx, int *ny,
                          int *nz, int *degree);

/* Variable-coefficient operator A(rho): init_ppe_proj_coefficients
 * (projection_method.cpp:404-465), face coefficient 1/(0.5*(rho_i+rho_j))/h^2,
 * zero at walls, diagonal -sum.  rho: n cells, MG layout, host pointer.
 * Coarse levels average the fine face coefficients on each coarse face.
 * Only valid with SCIHPC_MG_NEUMANN walls (as the reference's PPE). */
SCIHPC_API int scihpc_mg_set_density(scihpc_mg *h, const double *rho);
/* same, rho given as the padded level-set-grid density array
 * (pressure->dummy->density, projection_method.cpp:416-419) */
SCIHPC_API int scihpc_mg_set_density_padded(scihpc_mg *h, const double *rho_padded);

/* multigrid_base::rhs / sol / res of level `level` (multigrid_base.h:28):
 * host copies in MG layout, n = nx*ny*nz of that level. */
SCIHPC_API int scihpc_mg_set_rhs(scihpc_mg *h, int level, const double *src);
SCIHPC_API int scihpc_mg_get_rhs(scihpc_mg *h, int level, double *dst);
SCIHPC_API int scihpc_mg_set_sol(scihpc_mg *h, int level, const double *src);
SCIHPC_API int scihpc_mg_get_sol(scihpc_mg *h, int level, double *dst);
SCIHPC_API int scihpc_mg_get_res(scihpc_mg *h, int level, double *dst);
/* device-pointer variants for level 0 (d2d copies on the handle stream;
 * set_rhs_device returns without waiting: the copy is ordered before the
 * handle's next work, and the source must stay valid until it has run) */
SCIHPC_API int scihpc_mg_set_rhs_device(scihpc_mg *h, const double *dsrc);
SCIHPC_API int scihpc_mg_get_sol_device(scihpc_mg *h, double *ddst);
/* multigrid_base::no_compatibility (multigrid_base.h:33) */
SCIHPC_API int scihpc_mg_set_nocompat(scihpc_mg *h, int level, int flag);

/* multigrid_base::residual() (multigrid_base.cpp:69-72): res = rhs - A sol,
 * returns ||res||_2 (absolute) */
SCIHPC_API int scihpc_mg_residual(scihpc_mg *h, int level, double *norm);
/* No counterpart in the reference: a variant of scihpc_mg_residual for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No handle and no
 * allocation: res = b - A x on the level-0 grid of caller device arrays (MG
 * layout) and *norm = ||res||_2, both written to device memory, enqueued on
 * `stream` (a hipStream_t; NULL = the default stream) without a host sync.
 * A = A(rho) (face 1/(0.5*(rho_i+rho_j))/h^2, zero at walls, as
 * scihpc_mg_set_density; bc must be SCIHPC_MG_NEUMANN) when rho is given,
 * else the constant operator of `bc`.  res may be NULL.  The per-block
 * partial sums live in one fixed device array per GPU (512.5 KiB), which
 * limits the grid to 2^24 cells (256^3); larger grids return SCIHPC_BAD_ARG
 * before any launch.  Calls that overlap on one GPU race on that array. */
SCIHPC_API int scihpc_mg_residual_noalloc(int ndim, int nx, int ny, int nz, double dx, double dy,
                                          double dz, int bc, const double *rho, const double *x,
                                          const double *b, double *res, double *norm, void *stream);
/* multigrid_base::Ax (multigrid_base.cpp:388-390) on a host vector */
SCIHPC_API int scihpc_mg_apply(scihpc_mg *h, int level, const double *x, double *y);
/* multigrid_base::compatibility_condition (multigrid_base.cpp:47-67) on
 * which = 0 rhs, 1 sol */
SCIHPC_API int scihpc_mg_compat(scihpc_mg *h, int level, int which);

/* Reference-faithful pieces on the GPU (same algorithm as the reference):
 * relax = steepest descent (multigrid_base.cpp:16-45); restriction =
 * relax(10) + residual + child average (:74-113); prolongation = slope
 * reconstruction + relax(10) (:115-208); coarse_solve = direct solve with
 * refinement (:210-234); vcycle_ref / fullcycle_ref = multigrid.cpp:43-79. */
SCIHPC_API int scihpc_mg_relax(scihpc_mg *h, int level, int iters);
SCIHPC_API int scihpc_mg_restriction(scihpc_mg *h, int level);
SCIHPC_API int scihpc_mg_prolongation(scihpc_mg *h, int level);
SCIHPC_API int scihpc_mg_coarse_solve(scihpc_mg *h);
SCIHPC_API int scihpc_mg_vcycle_ref(scihpc_mg *h);
SCIHPC_API int scihpc_mg_fullcycle_ref(scihpc_mg *h);

/* The MI355X V-cycle (replaces multigrid::v_cycle, multigrid.cpp:43-55, as
 * the solver behind the same interface): FP64 red-black Gauss-Seidel
 * V(nu1,nu2), fused residual + child-average restriction, (bi/tri)linear
 * prolongation, dense coarsest solve; in place on level 0 (rhs, sol). */
SCIHPC_API int scihpc_mg_vcycle(scihpc_mg *h, int nu1, int nu2);

typedef struct {
  double rel_tol;    /* stop when ||b - A p||_2 / ||b||_2 <= rel_tol */
  int max_cycles;    /* cap on V-cycles (outer corrections in mixed mode) */
  int nu1, nu2;      /* smoothing sweeps down / up */
  int precision;     /* SCIHPC_PREC_FP64 or SCIHPC_PREC_MIXED */
  int use_learned;   /* 1: learned initial guess (if attached), judged by the
                        FP64 residual, with back-off: after the k-th
                        rejection in a row the next 2^k - 1 solves (at most
                        63) skip the network; 3: judged on every solve;
                        2: the guess kept unjudged (ablation) */
  int zero_guess;    /* 1: start from p = 0, else from the current level-0 sol */
  int accel;         /* 0: stationary defect correction (one V-cycle per
                        correction); 1: flexible CG (FCG(1)) with the V-cycle
                        as preconditioner, FP64 vectors and inner products */
  int learned_correction; /* mixed flexible-CG solves with a network attached:
                        1: after the first iteration the network maps the
                        current FP64 defect r (its block means on the
                        network's level, with ln rho) to a coarse-level
                        correction c = alpha S P S N(r) (alpha energy-
                        optimal); x += c is kept only if it lowers the FP64
                        residual, and then the CG recurrence restarts from
                        the corrected x; 2: kept unjudged (ablation);
                        0: off.  No counterpart in the reference. */
  int lagged_check;  /* 1: the host reads the residual only from one
                        iteration before the count that the last converged
                        solve on this handle (same path, tolerance and
                        nu1/nu2) needed, and on the last allowed iteration;
                        earlier iterations are enqueued without a host round
                        trip.  A solve may then run past the tolerance by a
                        few iterations (the count self-corrects); the result
                        still meets rel_tol.  0: read after every
                        iteration, as the reference's loop does.  No
                        counterpart in the reference. */
} scihpc_solve_opts;

typedef struct {
  int cycles;            /* V-cycles (= outer corrections in mixed mode) */
  int learned_accepted;  /* learned steps accepted by the FP64 residual test */
  int learned_rejected;  /* learned steps discarded (plain MG used instead) */
  double rel_residual;   /* final ||b - A p|| / ||b||, FP64 */
  double ms;             /* device time of the solve (HIP events) */
  double learned_residual; /* ||b - A p0|| / ||b|| of the learned guess (0 if none) */
  int learned_skipped;     /* 1: the back-off skipped the network on this solve */
  int correction_accepted; /* learned coarse-level corrections kept (learned_correction) */
  int correction_rejected; /* ... discarded by the FP64 residual test */
  double correction_residual; /* ||r - alpha A c|| / ||r|| of the correction (0 if none) */
} scihpc_solve_stats;

/* Solve A p = b on level 0 to opts->rel_tol (the call behind
 * projection_method::fast_pressure_correction's loop,
 * projection_method.cpp:694-706).  For Neumann operators b is used as given
 * (callers remove its mean, as the reference's drivers do) and p is returned
 * mean-free (compatibility_condition).  Returns SCIHPC_NOT_CONVERGED when
 * max_cycles is reached first. */
SCIHPC_API int scihpc_mg_solve(scihpc_mg *h, const scihpc_solve_opts *opts,
                    scihpc_solve_stats *stats);

/* Timing hook (no reference counterpart): runs `iters` launches of one
 * level-0 kernel between HIP events on the handle stream; which = 0 FP64
 * pre-smoothing (nu1 = 2 sweeps + residual + restriction), 1 the same in
 * FP32, 2 FP32 post-smoothing fused with the FP64 outer update, 3 FP64
 * post-smoothing with the residual norm, 4 the FP64 prolongation of level 1
 * added to level 0 alone, 5 the FP64 level-0 residual pass (partial sums of
 * the norm) alone, 6 the mixed flexible-CG update fused into the FP32
 * level-0 pre-smoothing (2-D; after a mixed flexible-CG solve, on its
 * vectors).  Returns the mean milliseconds per launch and the algorithmic
 * HBM bytes of one launch. */
SCIHPC_API int scihpc_mg_time_kernel(scihpc_mg *h, int which, int iters, double *ms,
                          double *bytes);

/* Measurement hook (no reference counterpart): HIP events on the handle
 * stream around every launch of one kernel inside the following solves, at
 * most `capacity` launches; which = 6 the mixed flexible-CG update fused
 * into the FP32 level-0 pre-smoothing (2-D), -1 disarms. */
SCIHPC_API int scihpc_mg_probe_kernel(scihpc_mg *h, int which, int capacity);
/* launches recorded since the probe was armed or last read; total_ms: the
 * summed time between the HIP events around them (each interval also holds
 * the dispatch of the launch); kernel_ms: the summed spans of the kernels
 * themselves, earliest workgroup start to latest workgroup end on the
 * device's wall clock; bytes: the algorithmic HBM bytes of one launch (as
 * scihpc_mg_time_kernel).  Starts a new count. */
SCIHPC_API int scihpc_mg_probe_result(scihpc_mg *h, int *launches, double *total_ms, double *kernel_ms,
                                      double *bytes);

/* ------------------------------------------------------------------ */
/* Projection step on padded fields (scalar_data layout), around a
 * multigrid handle of the same interior grid.  Velocities are the MAC face
 * components u, v, w; their BCs are applied after the correction as
 * wrapper::apply_vel_bc (wrapper.cpp:21-26, SLIP made NO_SLIP for the
 * wall-normal component, wrapper.cpp:111-125); the pressure gets Neumann
 * BCs.  Host arrays; calls synchronise.                                 */
typedef struct scihpc_proj scihpc_proj;
SCIHPC_API int scihpc_proj_create(int ndim, int nx, int ny, int nz, double dx, double dy,
                                  double dz, const int vel_bc_types[6], scihpc_proj **out);
SCIHPC_API int scihpc_proj_destroy(scihpc_proj *h);

typedef struct {
  double dt, dt_old, dt_old2; /* problem_parameters (global.h:25-54) */
  double rho12;               /* density_ratio */
  double ppe_tol;             /* absolute residual of the multigrid loop */
  int max_cycles;             /* the reference caps the loop at 100 */
  int cycle;                  /* 0: the reference's full cycles; 1: this
                                 build's red-black V(2,2) (scihpc_mg_vcycle),
                                 same loop and stopping rule */
} scihpc_fpc_params;

/* One FPC step: pressure_guess_solution (projection_method.cpp:803-815),
 * init_ppe_fpc_coefficients (:467-541) and fast_pressure_correction
 * (:659-753): p_hat from p_old, p_old2; MG rhs = rho12/dt div u + sum C
 * p_hat; full cycles of the reference-faithful multigrid (a Neumann
 * constant-coefficient handle whose level-0 sol is the warm start, as
 * SciHPC's persistent at[0]->sol) until ||r|| <= ppe_tol (with cycle = 1
 * red-black V(2,2) cycles instead, no counterpart in the reference); p,
 * velocity correction, BCs.  Writes u, v, w, p, p_hat and the cycle
 * count. */
SCIHPC_API int scihpc_proj_fpc_step(scihpc_proj *h, scihpc_mg *mg, const scihpc_fpc_params *prm,
                                    const double *rho, const double *p_old,
                                    const double *p_old2, double *u, double *v, double *w,
                                    double *p, double *phat, int *cycles);

/* The variable-density projection (projection_method.cpp:544-657) with the
 * SOR replaced by scihpc_mg_solve on A(rho): rhs = div u / dt (mean
 * removed), A(rho) p = rhs, u -= dt / (rho_face h) dp, BCs.  `mg` must be a
 * Neumann handle; rho is padded. */
SCIHPC_API int scihpc_proj_step(scihpc_proj *h, scihpc_mg *mg, double dt,
                                const scihpc_solve_opts *opts, const double *rho, double *u,
                                double *v, double *w, double *p, scihpc_solve_stats *stats);

/* Momentum predictor of ab_solve_sec (projection_method.cpp:779-783).
 * Flow parameters of problem_parameters (global.h:25-54); froude <= 0
 * disables gravity and weber <= 0 surface tension, as in the reference. */
typedef struct {
  double dt;
  double reynolds, froude, weber;
  double ls_width;                       /* interface half width (1.5 h in the drivers) */
  double density_ratio, viscosity_ratio; /* second fluid / first fluid */
  int positive_ref;                      /* 1: phi > 0 is the reference fluid */
} scihpc_flow_params;

/* projection_method::find_source_sec (:314-382) + find_intermediate_velocity
 * (:384-402): Heaviside density/viscosity and delta of phi, second-order
 * level-set derivatives, per direction the face transfer all_to_face_*
 * (wrapper_func.cpp:177-262) into nvel with its BCs, convection_sec
 * (source.cpp:42-73) and add_stress_*, gravity, surface tension (curvature,
 * wrapper_func.cpp:115-172), then AB2 u += dt (1.5 s - 0.5 s_old) and the
 * velocity BCs.  The sources u_src, v_src, w_src are handle state (zero
 * after create, as the reference's arrays; see set/get_sources).  phi, u,
 * v, w, nu, nv, nw padded host arrays; u, v, w, nu, nv, nw are updated;
 * rho (nullable) receives find_density (the pressure step's density).
 * Velocity BCs: the handle's (both for vel and nvel, as the drivers). */
SCIHPC_API int scihpc_proj_momentum_step(scihpc_proj *h, const scihpc_flow_params *prm,
                                         const double *phi, double *u, double *v, double *w,
                                         double *nu, double *nv, double *nw, double *rho);
/* projection_method::find_source_sec (projection_method.cpp:314-382)
 * alone: the sources of scihpc_proj_momentum_step without the velocity
 * update.  The handle's sources rotate (the previous ones are kept for the
 * AB2 step, as u_src_old); u, v, w are read only; nu, nv, nw receive the
 * last face transfer; rho (nullable) find_density. */
SCIHPC_API int scihpc_proj_find_sources(scihpc_proj *h, const scihpc_flow_params *prm, const double *phi,
                                        const double *u, const double *v, const double *w, double *nu,
                                        double *nv, double *nw, double *rho);
/* projection_method::find_intermediate_velocity (:384-402): u += dt (1.5
 * u_src - 0.5 u_src_old) (v, w alike) with the handle's sources, then the
 * velocity BCs. */
SCIHPC_API int scihpc_proj_intermediate_velocity(scihpc_proj *h, const scihpc_flow_params *prm, double *u,
                                                 double *v, double *w);
/* projection_method::u_src/v_src/w_src (projection_method.h:18) */
SCIHPC_API int scihpc_proj_set_sources(scihpc_proj *h, const double *su, const double *sv,
                                       const double *sw);
SCIHPC_API int scihpc_proj_get_sources(scihpc_proj *h, double *su, double *sv, double *sw);
/* node_from_face + nvel->apply_nvel_bc (wrapper_func.cpp:264-293,
 * wrapper.cpp:28-33): cell-centred velocities after the pressure step */
SCIHPC_API int scihpc_proj_node_velocity(scihpc_proj *h, const double *u, const double *v,
                                         const double *w, double *nu, double *nv, double *nw);
/* Times the device part of scihpc_proj_momentum_step on the resident
 * fields: average ms per step over `iters`. */
SCIHPC_API int scihpc_proj_time_momentum(scihpc_proj *h, const scihpc_flow_params *prm, int iters,
                                         double *ms);

/* ------------------------------------------------------------------ */
/* Learned pressure guess (U-Net).  No reference counterpart: SciHPC has no
 * learned component.  A 4-level U-Net (16/32/64/128 channels, 3x3 (2-D) or
 * 3x3x3 (3-D) convolutions, BF16 weights and activations, FP32 accumulation
 * on the matrix cores) maps (b / sigma_b, ln rho / ln 1000) on one
 * multigrid level to p / (sigma_b * out_scale) there; the solve prolongs,
 * smooths and scales it and keeps it only if it lowers the FP64 residual.
 * Pixels are x fastest: (D x) H x W.                                      */
typedef struct scihpc_unet scihpc_unet;

/* 2-D network: height x width must be multiples of 8; max_batch images per
 * pass */
SCIHPC_API int scihpc_unet_create(int height, int width, int max_batch,
                                  unsigned long long seed, scihpc_unet **out);
/* 3-D network (3x3x3 convolutions, 2x2x2 pooling): depth x height x width,
 * each a multiple of 8 */
SCIHPC_API int scihpc_unet_create3(int depth, int height, int width, int max_batch,
                                   unsigned long long seed, scihpc_unet **out);
SCIHPC_API int scihpc_unet_destroy(scihpc_unet *h);
SCIHPC_API int scihpc_unet_num_params(scihpc_unet *h, long long *n);
/* layer 0..13: convolutions (w [cout][3][3](3)[cin] then b [cout] at
 * offset); layer 14: the 1x1 head (16 weights then 1 bias) */
SCIHPC_API int scihpc_unet_layer(scihpc_unet *h, int layer, int *cin, int *cout,
                                 int *level, long long *offset);
SCIHPC_API int scihpc_unet_get_params(scihpc_unet *h, float *host);
SCIHPC_API int scihpc_unet_set_params(scihpc_unet *h, const float *host);
SCIHPC_API int scihpc_unet_get_grads(scihpc_unet *h, float *host);
SCIHPC_API int scihpc_unet_set_out_scale(scihpc_unet *h, float scale);
SCIHPC_API int scihpc_unet_get_out_scale(scihpc_unet *h, float *scale);
/* training samples kept on the device: feat [n][pixels][2], target
 * [n][pixels], faces [n][pixels][ndim] (host FP32, as
 * scihpc_mg_learned_features writes them).
 * With faces the loss is the A(rho) energy norm of the error relative to the
 * target's, sum_f c_f (e_i - e_j)^2 / sum_f c_f (t_i - t_j)^2 (what flexible
 * CG reduces); with faces = NULL it is the mean squared error.  Set
 * out_scale first: targets are divided by it. */
SCIHPC_API int scihpc_unet_set_samples(scihpc_unet *h, int n, const float *feat,
                                       const float *target, const float *faces);
/* forward pass on stored samples; idx [n][2] = (sample, flips: bit 0 mirrors
 * x, bit 1 mirrors y, bit 2 mirrors z); out (host [n][pixels]) may be NULL */
SCIHPC_API int scihpc_unet_forward(scihpc_unet *h, int n, const int *idx, float *out);
/* loss of the last forward pass against target / out_scale, and all
 * parameter gradients (read with scihpc_unet_get_grads) */
SCIHPC_API int scihpc_unet_backward(scihpc_unet *h, double *loss);
/* forward + backward + one Adam step (lr); loss may be NULL (no sync) */
SCIHPC_API int scihpc_unet_train_step(scihpc_unet *h, int n, const int *idx, float lr,
                                      double *loss);
/* energy loss of the smoothed error: E(M^sweeps e) / E(t) with M one
 * damped-Jacobi sweep (weight omega in (0, 1]) of the network grid's A(rho),
 * the error the multigrid smoother leaves of the network's output (0 sweeps:
 * the plain energy norm, the default) */
SCIHPC_API int scihpc_unet_set_loss_smoothing(scihpc_unet *h, int sweeps, float omega);
SCIHPC_API int scihpc_unet_sync(scihpc_unet *h);
/* device ms of one forward pass (which 0) or training step (which 1, with
 * a negligible learning rate) on stored samples 0..n-1, and its MFMA FLOPs */
SCIHPC_API int scihpc_unet_time(scihpc_unet *h, int n, int which, int iters, double *ms,
                                double *flops);
/* single 3x3 convolution (NHWC BF16 x [n][h][w][cin], w [cout][3][3][cin]
 * BF16, b [cout] FP32 or NULL, y [n][h][w][cout]; host arrays, copied to
 * and from the device); cin a multiple of 16, cout = 16 x {1,2,3,4,6,8,12} */
SCIHPC_API int scihpc_conv3x3_forward(const uint16_t *x, int n, int height, int width,
                                      int cin, const uint16_t *w, const float *b,
                                      int cout, int relu, uint16_t *y);
/* single 3x3x3 convolution (NDHWC BF16 x [n][d][h][w][cin], w
 * [cout][3][3][3][cin]), otherwise as scihpc_conv3x3_forward */
SCIHPC_API int scihpc_conv3x3x3_forward(const uint16_t *x, int n, int depth, int height,
                                        int width, int cin, const uint16_t *w,
                                        const float *b, int cout, int relu, uint16_t *y);

/* Use `net` (NULL detaches) for the learned initial guess of
 * scihpc_mg_solve(use_learned = 1); its size must equal multigrid level
 * `level`'s grid.  2-D only.  The guess (network output, smoothed and
 * prolonged level by level, scaled by the energy-optimal factor) is judged
 * by its FP64 residual on every solve.  In the mixed flexible-CG solve it is
 * formed on a second stream while the iterations from p = 0 run, and when it
 * is ready it replaces the iterate only if its residual is below both ||b||
 * and the iterate's; otherwise it is formed first and kept only if its
 * residual is below ||b||.  A rejected guess leaves plain multigrid from
 * p = 0 (learned_rejected). */
SCIHPC_API int scihpc_mg_attach_learned(scihpc_mg *h, scihpc_unet *net, int level);
/* The network's input features of the current level-0 b (and density) on
 * `level`: feat [H*W][2]; target [H*W] (may be NULL) = block mean of the
 * current level-0 solution / sigma_b, the training target; faces [H*W][2]
 * (may be NULL) = the level's face coefficients x h^2 toward x+1 and y+1. */
SCIHPC_API int scihpc_mg_learned_features(scihpc_mg *h, int level, float *feat,
                                          float *target, float *faces);

/* ------------------------------------------------------------------ */
/* Level-set advection: WENO5-Z derivatives + TVD-RK3 (+ BCs)          */
typedef struct scihpc_ls scihpc_ls;

/* weno_solver(scalar_data*, structured_grid*) (weno_solver.cpp:7-14) and
 * runge_kutta(Nx,Ny,Nz) (runge_kutta.cpp:7-11) for one padded grid.
 * bc_types/bc_values: 6 entries xl,xr,yl,yr,zl,zr (bc_factory.h:42-48). */
SCIHPC_API int scihpc_ls_create(int ndim, int nx, int ny, int nz, double dx, double dy,
                     double dz, const int *bc_types, const double *bc_values,
                     scihpc_ls **out);
SCIHPC_API int scihpc_ls_destroy(scihpc_ls *h);
/* padded host arrays: phi (scalar->data), u/v/w node velocities (nvel) */
SCIHPC_API int scihpc_ls_set_phi(scihpc_ls *h, const double *phi);
SCIHPC_API int scihpc_ls_get_phi(scihpc_ls *h, double *phi);
SCIHPC_API int scihpc_ls_set_velocity(scihpc_ls *h, const double *u, const double *v,
                           const double *w);
/* weno_solver::weno5_find_derivatives (weno_solver.cpp:270-278) of `flux`
 * (padded, host); fx/fy/fz padded outputs, interior cells written. */
SCIHPC_API int scihpc_ls_weno5_derivatives(scihpc_ls *h, const double *flux, double *fx,
                                double *fy, double *fz);
/* runge_kutta::tvd_rk3(phi, nvel, identity_flux, convection) with WENO
 * derivatives (runge_kutta.cpp:13-49, source.cpp:8-40, flux.cpp:7-17),
 * bc_factory::apply_node_centered after each stage; nsteps steps of dt. */
SCIHPC_API int scihpc_ls_tvd_rk3(scihpc_ls *h, double dt, int nsteps);
/* bc_factory::apply_node_centered / apply_node_staggered_{x,y,z}
 * (bc_factory.cpp:61-375) on a padded host array with this handle's BCs;
 * stag = -1 node centred, 0/1/2 staggered axis. */
SCIHPC_API int scihpc_ls_apply_bc(scihpc_ls *h, double *f, int stag);
/* find_density (wrapper_func.cpp:387-402, lsm.cpp:8-18) of the handle's
 * phi into a padded host array. */
SCIHPC_API int scihpc_ls_density(scihpc_ls *h, double ls_width, double ratio,
                      int positive_ref, double *rho);
/* Timing hook: `iters` TVD-RK3 steps between HIP events; mean ms per step
 * and algorithmic HBM bytes per step. */
SCIHPC_API int scihpc_ls_time_step(scihpc_ls *h, double dt, int iters, double *ms,
                        double *bytes);

/* Interface maintenance (the level set between flow steps). */
enum { SCIHPC_LS_WENO5 = 0, SCIHPC_LS_UCCD = 1 };
/* Convection scheme of scihpc_ls_tvd_rk3: WENO5 (weno_solver.cpp, the
 * default) or UCCD, the reference's own convection (source.cpp:8-40 with
 * uccd_solver.cpp:14-215: compact upwind/downwind line systems chosen by the
 * sign of the node velocity), updated on every padded cell as the reference. */
SCIHPC_API int scihpc_ls_set_scheme(scihpc_ls *h, int scheme);
/* How the CCD/UCCD line systems are solved (the twin back substitution of
 * matrix_solver.cpp:58-83 on the factorised coefficients of twin_dec,
 * :21-56).  partitioned = 0 (the default): one thread per line in the
 * reference's order, bit-identical to it.  partitioned = 1: every line cut
 * into segments that sweep in parallel and exchange their incoming states
 * through 2x2 segment propagators, then sweep again from the true states
 * with the reference's arithmetic; the results differ from the sequential
 * sweep only by the rounding of the carried states, within the sequential
 * sweep's own rounding error (lines longer than 2176 points always sweep
 * sequentially).  No counterpart in the reference (its solve is always
 * sequential); for callers that trade the reference's bits for speed. */
SCIHPC_API int scihpc_ls_set_line_solver(scihpc_ls *h, int partitioned);
/* uccd_solver::find_derivatives(f, vel) (upwind = 1, the handle's
 * velocity) or ccd_solver::find_derivatives(f) (ccd_solver.cpp:14-142) of
 * a padded host array: first and second derivatives along every axis. */
SCIHPC_API int scihpc_ls_compact_derivatives(scihpc_ls *h, const double *f, int upwind, double *fx,
                                             double *fy, double *fz, double *fxx, double *fyy,
                                             double *fzz);
/* lsf_mass (wrapper_func.cpp:450-511) of the current phi (2-D, 3-D). */
SCIHPC_API int scihpc_ls_mass(scihpc_ls *h, double ls_width, double ratio, int positive_ref,
                              double *mass);
/* `count` runge_kutta::euler steps with mpls (source.cpp:109-146): phi
 * moved along its normal by eta delta |grad phi| so that lsf_mass returns
 * toward mass0 (params->lsf_mass0); BCs after each step.  Leaves delta of
 * the last pre-update phi (the reference's dummy->delta). */
SCIHPC_API int scihpc_ls_mpls(scihpc_ls *h, double dt, double ls_width, double ratio,
                              int positive_ref, double mass0, int count);
/* find_sign / find_delta (wrapper_func.cpp:20-40) of the current phi into
 * the handle's sign / delta (the dummy arrays the redistancing reads). */
SCIHPC_API int scihpc_ls_find_sign(scihpc_ls *h, double ls_width);
SCIHPC_API int scihpc_ls_find_delta(scihpc_ls *h, double ls_width);
enum { SCIHPC_LS_SIGN = 0, SCIHPC_LS_DELTA = 1 };
SCIHPC_API int scihpc_ls_set_field(scihpc_ls *h, int which, const double *src);
SCIHPC_API int scihpc_ls_get_field(scihpc_ls *h, int which, double *dst);
/* `nsteps` runge_kutta::tvd_rk3 steps with lsf_redistance_lambda
 * (source.cpp:148-240, godunov_gradient.cpp:7-201): pseudo-time rdt,
 * the stored sign and delta (not recomputed, as in the reference). */
SCIHPC_API int scihpc_ls_redistance(scihpc_ls *h, double dt, double rdt, int nsteps);
/* stabilized_upon_gradient (godunov_gradient.cpp:159-181): the Godunov
 * gradient of phi with the stored sign (godunov_gradient, :7-152) on every
 * padded cell, phi /= its maximum.  grad (nullable, padded host array)
 * receives the gradient (the reference's dummy->grad), max_grad (nullable)
 * the maximum. */
SCIHPC_API int scihpc_ls_stabilize_gradient(scihpc_ls *h, double *grad, double *max_grad);

/* ------------------------------------------------------------------ */
/* Snapshots: vtkWriter (vtkWriter.h:15-35, vtkWriter.cpp:18-111).  Files are
 * byte-identical to the reference's legacy VTK (ASCII header, BINARY
 * STRUCTURED_POINTS, CELL_DATA, big-endian FP64, x fastest).  Fields are
 * padded (scalar_data layout) host arrays (on_device = 0) or device arrays
 * (on_device = 1, complete before the call; the device does the reorder
 * and byte swap).  add_* return once the field has been read; the file is
 * written by a background thread and is complete after scihpc_vtk_close
 * (or the next scihpc_vtk_begin), which also report write failures as
 * SCIHPC_IO_ERROR.                                                       */
typedef struct scihpc_vtk scihpc_vtk;
/* vtkWriter(structured_grid*, filename): origin = axis starts, spacing =
 * dx, dy, dz of the grid (entries beyond ndim are ignored and written as 0,
 * as structured_grid.cpp:7-84 sets them). */
SCIHPC_API int scihpc_vtk_create(int ndim, int nx, int ny, int nz, const double origin[3],
                                 const double spacing[3], const char *prefix, scihpc_vtk **out);
SCIHPC_API int scihpc_vtk_destroy(scihpc_vtk *h);
/* vtkWriter::create(id): opens prefix<id>.vtk and writes the header */
SCIHPC_API int scihpc_vtk_begin(scihpc_vtk *h, unsigned int id);
/* vtkWriter::add_scalar(scalar_data*, name) / add_scalar(DataType***, ...) */
SCIHPC_API int scihpc_vtk_add_scalar(scihpc_vtk *h, const char *name, const double *field,
                                     int on_device);
/* vtkWriter::add_vector(vector_data*, name) / add_vector(u, v, w, ...) */
SCIHPC_API int scihpc_vtk_add_vector(scihpc_vtk *h, const char *name, const double *u,
                                     const double *v, const double *w, int on_device);
/* vtkWriter::close(): waits for the writer thread and closes the file */
SCIHPC_API int scihpc_vtk_close(scihpc_vtk *h);
/* Times the reorder + byte-swap kernel alone (ncomp 1: a scalar, 3: a
 * vector) on the handle's own device staging fields: average ms per launch
 * over `iters`.  Algorithmic bytes per launch: 16 * ncomp * cells. */
SCIHPC_API int scihpc_vtk_time_pack(scihpc_vtk *h, int ncomp, int iters, double *ms);

#ifdef __cplusplus
}
#endif
#endif
