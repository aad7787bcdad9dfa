/*
 * oracle.h -- CPU restatement of SciHPC's projection-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in scihpc_amd/ links, loads or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the SciHPC source tree).  Pinning status of each function (what checks it
 * besides retyping its formula) is listed next to its declaration:
 *   [golden]  compared with vectors written by the compiled reference
 *   [math]    closed form / invariant / brute force checks in tests/
 *   [unpinned] nothing independent checks it yet
 *
 * Layouts
 *   MG vector : unpadded, x fastest, idx = i + nx*(j + ny*k)
 *               (multigrid_base.cpp:8-14)
 *   padded    : reference scalar_data layout data[i][j][k] made contiguous,
 *               k fastest, idx = (I*Ny + J)*Nz + K with ghost width ghc=3 on
 *               every axis whose interior size is > 1 (scalar_data.cpp:73-158)
 *
 * BC type codes: 0 neumann, 1 dirichlet, 2 periodic, 3 no_slip, 4 slip
 * (bc_factory.h:18-22).  Face order in 6-arrays: xl, xr, yl, yr, zl, zr.
 */
#ifndef SCIHPC_ORACLE_H
#define SCIHPC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* OpenMP threads of the loops that write disjoint cells (every reduction
 * stays serial, so results do not depend on n); returns the count in use. */
int orc_set_threads(int n);

/* ------------------------------------------------------------------ */
/* multigrid hierarchy (multigrid.cpp:8-41, 81-122)          [golden]   */
int orc_mg_levels(int ndim, int nx, int ny, int nz, int maxl,
                  int *lnx, int *lny, int *lnz, int *ldeg);

/* ------------------------------------------------------------------ */
/* multigrid object.  bc: 0 = init_NeumannBC (multigrid_base.cpp:236-291),
 * 1 = init_DirichletBC (multigrid_base.cpp:293-345).  Level spacing is
 * dx*degree (multigrid.cpp:116-119).                                  */
void *orc_mg_create(int ndim, int nx, int ny, int nz,
                    double dx, double dy, double dz, int bc);
void orc_mg_destroy(void *h);
int orc_mg_nlevels(void *h);
void orc_mg_level_shape(void *h, int level, int *shape /* nx,ny,nz,deg,n */);
/* which: 0 rhs, 1 sol, 2 res */
double *orc_mg_vec(void *h, int level, int which);
void orc_mg_set_nocompat(void *h, int level, int flag);

/* Variable coefficient operator A(rho) on level 0, face coefficient
 * 1/(0.5*(rho_i+rho_nb))/dx/dx, zero on wall faces, diagonal
 * -(CR+CL+CU+CD)-(CF+CB) (projection_method.cpp:404-465).  Coarse levels
 * get face coefficients averaged over the fine faces lying on each coarse
 * face, scaled by (h_fine/h_coarse)^2 (build's choice, no reference).
 * rho: MG layout, level-0 cells.                           [golden lvl0] */
void orc_mg_set_rho(void *h, const double *rho);
/* read back the level's operator: diag and face arrays (cx = face i+1/2) */
void orc_mg_get_operator(void *h, int level, double *diag, double *cx,
                         double *cy, double *cz);

/* reference-faithful pieces (multigrid_base.cpp)                      */
void orc_mg_apply(void *h, int level, const double *x, double *y); /* :388-390 [golden] */
double orc_mg_residual(void *h, int level);        /* :69-72   [golden] */
void orc_mg_compat(void *h, int level, int which); /* :47-67   [math]   */
void orc_mg_relax(void *h, int level, int iter);   /* :16-45   [golden] */
void orc_mg_restriction(void *h, int level);       /* :74-113  [golden] */
void orc_mg_prolongation(void *h, int level);      /* :115-208 [golden] */
void orc_mg_coarse_solve(void *h);                 /* :210-234 [golden] */
void orc_mg_vcycle_ref(void *h);                   /* multigrid.cpp:43-55 [golden] */
void orc_mg_fullcycle_ref(void *h);                /* multigrid.cpp:57-79 [golden] */

/* build's algorithm (no reference counterpart; pinned by converging to the
 * same discrete solution as the reference cycles and a dense solve)  [math]
 * V(nu1,nu2): red-black Gauss-Seidel (red first going down, black first
 * going up), residual, child-average restriction, cell-centred
 * (bi/tri)linear prolongation, dense coarsest solve (pinned for Neumann). */
/* transfer kinds: rkind 0 child average (default), 1 full weighting
 * (P^T/r^d); pkind 0 linear (default), 1 piecewise constant           */
void orc_mg_set_transfer(void *h, int rkind, int pkind);
void orc_mg_smooth_rb(void *h, int level, int sweeps, int black_first);
void orc_mg_restrict_rb(void *h, int level);
void orc_mg_prolong_rb(void *h, int level);
void orc_mg_coarse_direct(void *h);
void orc_mg_vcycle_rb(void *h, int nu1, int nu2);
/* V-cycles on level 0 until ||rhs-A sol||/||rhs|| <= rel_tol; returns cycles
 * used, writes the relative residual after each cycle into hist[0..]    */
int orc_mg_solve_rb(void *h, double rel_tol, int max_cycles, int nu1, int nu2,
                    double *hist);
/* flexible CG, FCG(1) (Notay 2000), with one V(nu1,nu2) of
 * orc_mg_vcycle_rb from zero as the preconditioner; b = level 0's rhs, x0 =
 * 0, stops when the recurrence residual ||r||/||b|| <= rel_tol and the true
 * one b - A x agrees (else it restarts from x); on all-Neumann operators
 * the step sizes leave out the direction's constant component; result in
 * level 0's sol (mean-free for Neumann); hist[k] = relative residual after
 * iteration k.  Returns iterations.  Build's outer loop (closest analogue:
 * projection_method.cpp:694-706).                                [math]
 * Pinned by the dense solve, the A-conjugacy p.q_old = 0 and the
 * orthogonality p.r = 0 after each step, and the absence of drift below the
 * FP64 floor (tests/test_oracle_math.py). */
int orc_mg_solve_fcg(void *h, double rel_tol, int max_iter, int nu1, int nu2,
                     double *hist);
/* its level-0 vectors after the solve: 0 x (before the mean removal), 1 r,
 * 2 p, 3 q, 4 p_old, 5 q_old, 6 b - A x of the final x            */
double *orc_mg_fcg_vec(void *h, int which);
/* dense direct solve of level 0 (n <= 4096): Dirichlet -> plain LU;
 * Neumann -> pinned system, mean-free result.                   [math]   */
int orc_mg_dense_solve_level0(void *h, const double *b, double *x);

/* ------------------------------------------------------------------ */
/* padded-field kernels                                                */
int orc_padded_size(int ndim, int nx, int ny, int nz, int *N /*Nx,Ny,Nz*/);

/* bc_factory::apply_node_centered (bc_factory.cpp:61-138) when stag < 0,
 * apply_node_staggered_{x,y,z} (bc_factory.cpp:140-375) when stag = 0/1/2.
 * Per-type formulas: neumann.cpp, dirichlet.cpp, periodic.cpp, no_slip.cpp,
 * slip.cpp.                                                   [golden]  */
void orc_bc_apply(int ndim, int nx, int ny, int nz, double dx, double dy,
                  double dz, const int *types, const double *values,
                  int stag, double *f);

/* weno_solver::weno5_find_derivatives (weno_solver.cpp:16-278): writes
 * fx/fy/fz on interior cells only; fm/fp/fh are padded scratch. [golden] */
void orc_weno5_derivatives(int ndim, int nx, int ny, int nz, double dx,
                           double dy, double dz, const double *flux,
                           const double *u, const double *v, const double *w,
                           double *fx, double *fy, double *fz);

/* convection with WENO derivatives (source.cpp:8-40 with the weno call of
 * :11): identity_flux (flux.cpp:7-17), derivatives, s over the padded
 * array.                                                      [golden]  */
void orc_convection_weno(int ndim, int nx, int ny, int nz, double dx,
                         double dy, double dz, const double *phi,
                         const double *u, const double *v, const double *w,
                         double *fx, double *fy, double *fz, double *s);

/* runge_kutta::tvd_rk3 (runge_kutta.cpp:13-49) with rhs = convection_weno
 * and bc_factory::apply_node_centered after each stage.  fx/fy/fz persist
 * between stages exactly like scalar_data's arrays.            [golden]  */
void orc_tvd_rk3_weno(int ndim, int nx, int ny, int nz, double dx, double dy,
                      double dz, double dt, const int *types,
                      const double *values, double *phi, const double *u,
                      const double *v, const double *w, double *fx,
                      double *fy, double *fz, int nsteps);

/* lsm.cpp:8-18 Heaviside; wrapper_func.cpp:387-402 find_density [golden] */
void orc_density(long n, const double *phi, double ls_width, double ratio,
                 int positive_ref, double *rho);

/* projection_method.cpp:404-465 init_ppe_proj_coefficients on padded
 * arrays (interior cells written).                             [golden] */
void orc_ppe_coefficients(int ndim, int nx, int ny, int nz, double dx,
                          double dy, double dz, const double *rho, double *CC,
                          double *CR, double *CL, double *CU, double *CD,
                          double *CF, double *CB);
/* projection_method.cpp:546-568 RHS = div(u*)/dt on padded arrays [golden] */
void orc_projection_rhs(int ndim, int nx, int ny, int nz, double dx,
                        double dy, double dz, double dt, const double *u,
                        const double *v, const double *w, double *rhs);
/* projection_method.cpp:630-654 velocity update u -= dt/dx/rho_face*dp [golden] */
void orc_projection_velocity(int ndim, int nx, int ny, int nz, double dx,
                             double dy, double dz, double dt,
                             const double *rho, const double *p, double *u,
                             double *v, double *w);
/* projection_method.cpp:572-626 SOR sweeps on the padded pressure, run
 * serially (the reference loop is a data race under OpenMP).  Neumann BC
 * after each sweep as the pressure wrapper does.  Returns sweeps. [golden] */
int orc_projection_sor(int ndim, int nx, int ny, int nz, double omega,
                       int max_iter, double tol, double tol2,
                       const double *CC, const double *CR, const double *CL,
                       const double *CU, const double *CD, const double *CF,
                       const double *CB, const double *rhs, double *p,
                       double *tmp);

/* projection_method.cpp:467-541 FPC coefficients and RHS_fpc = sum C*p_hat
 * (p_hat = pressure->scalar->flux).                            [golden] */
void orc_fpc_coefficients(int ndim, int nx, int ny, int nz, double dx,
                          double dy, double dz, double rho12,
                          const double *rho, const double *phat, double *CC,
                          double *CR, double *CL, double *CU, double *CD,
                          double *CF, double *CB, double *RHS);
/* projection_method.cpp:667-692 MG rhs (x-fastest) = rho12/dt*div u + RHS
 *                                                 [golden via orc_fpc_step] */
void orc_fpc_rhs(int ndim, int nx, int ny, int nz, double dx, double dy,
                 double dz, double dt, double rho12, const double *u,
                 const double *v, const double *w, const double *RHS,
                 double *mg_rhs);
/* projection_method.cpp:708-720 p <- MG sol (padded <- x-fastest)        */
void orc_mg_to_padded(int ndim, int nx, int ny, int nz, const double *sol,
                      double *p);
void orc_padded_to_mg(int ndim, int nx, int ny, int nz, const double *p,
                      double *sol);
/* projection_method.cpp:723-750 FPC velocity correction [golden via orc_fpc_step] */
void orc_fpc_velocity(int ndim, int nx, int ny, int nz, double dx, double dy,
                      double dz, double dt, double rho12, const double *rho,
                      const double *p, const double *phat, double *u,
                      double *v, double *w);
/* projection_method.cpp:803-815 pressure guess p_hat  [golden via orc_fpc_step] */
void orc_pressure_guess(long n, double dt_old, double dt_old2,
                        const double *old, const double *old2, double *phat);
/* One FPC projection step: pressure_guess_solution, init_ppe_fpc_coefficients,
 * fast_pressure_correction (projection_method.cpp:467-541, 659-753, 803-815)
 * with the reference-faithful multigrid started from p = 0 (full cycles
 * until the absolute residual <= ppe_tol, at most max_cycles), Neumann
 * pressure BCs and staggered velocity BCs of vel_types (SLIP becomes
 * NO_SLIP for the wall-normal component, as the velocity wrapper sets it,
 * wrapper.cpp:111-125).  Returns cycles.
 *                                                              [golden] */
int orc_fpc_step(int ndim, int nx, int ny, int nz, double dx, double dy,
                 double dz, double dt, double dt_old, double dt_old2,
                 double rho12, double ppe_tol, int max_cycles,
                 const int *vel_types, const double *rho, const double *p_old,
                 const double *p_old2, double *u, double *v, double *w,
                 double *pr, double *phat);

/* ------------------------------------------------------------------ */
/* legacy VTK snapshot file: vtkWriter::create, add_scalar (nscalar
 * padded fields), add_vector (nvector padded u,v,w triples, vectors[3q+c]),
 * close (vtkWriter.cpp:18-111).  Returns 0, or -1 on an I/O failure.
 *                                                              [golden] */
int orc_vtk_write(const char *path, int ndim, int nx, int ny, int nz,
                  const double *origin, const double *spacing, int nscalar,
                  const char *const *snames, const double *const *scalars,
                  int nvector, const char *const *vnames,
                  const double *const *vectors);

/* ------------------------------------------------------------------ */
/* Momentum predictor: projection_method::find_source_sec
 * (projection_method.cpp:314-382: find_delta/density/viscosity, level-set
 * second-order derivatives, per direction all_to_face + nvel BCs +
 * convection_sec + add_stress, add_body_force, add_surface_force) and
 * find_intermediate_velocity (:384-402, AB2 + velocity BCs).
 * prm8 = {dt, Re, Fr, We, ls_width, density_ratio, viscosity_ratio,
 * positive_ref}.  vel* in/out; nv* (nvel) in/out; src* in = previous
 * sources (u_src, v_src, w_src), out = new sources; rho_out (nullable) =
 * find_density.  ndim 2 or 3.  Returns 0 (-1 if out of memory). [golden] */
int orc_momentum_step(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                      const double *prm8, const int *vel_types, const double *phi, double *vel0,
                      double *vel1, double *vel2, double *nv0, double *nv1, double *nv2,
                      double *src0, double *src1, double *src2, double *rho_out);
/* node_from_face + apply_nvel_bc (wrapper_func.cpp:264-293,
 * wrapper.cpp:28-33)                                           [golden] */
void orc_node_velocity(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                       const int *vel_types, const double *u, const double *v, const double *w,
                       double *nu, double *nv, double *nw);

/* ------------------------------------------------------------------ */
/* Interface maintenance (the level set between flow steps).
 * CCD / UCCD line derivatives (ccd_solver.cpp:14-142, uccd_solver.cpp:
 * 14-206, coefficients ccd_base.cpp:43-104 / uccd_base.cpp:16-117, twin
 * block solver matrix_solver.cpp:21-83) on the padded flux; upwind = 1
 * selects UCCD by the sign of u/v/w.                          [golden] */
void orc_compact_derivatives(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                             const double *flux, const double *u, const double *v, const double *w,
                             int upwind, double *fx, double *fy, double *fz, double *fxx, double *fyy,
                             double *fzz);
/* the same with the twin sweeps in long double (64-bit significand), rounded
 * to double at the end: the yardstick of a line solver's rounding error
 * [math: pinned by agreeing with the double sweep to within that sweep's
 * rounding error, and with the analytic derivatives of smooth fields] */
void orc_compact_derivatives_ld(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                                const double *flux, const double *u, const double *v, const double *w,
                                int upwind, double *fx, double *fy, double *fz, double *fxx, double *fyy,
                                double *fzz);
void orc_ccd_coefficients(int n, double h, double *c);
void orc_uccd_coefficients(int n, double h, double *up, double *dn);
/* tvd_rk3 with the UCCD convection term (source.cpp:8-40)     [golden] */
void orc_tvd_rk3_uccd(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                      const int *types, const double *values, double *phi, const double *u,
                      const double *v, const double *w, int nsteps);
/* the same on orc_compact_derivatives_ld's derivatives (the yardstick of a
 * line solver's rounding error after advection)
 * [math: pinned through orc_compact_derivatives_ld] */
void orc_tvd_rk3_uccd_ld(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                         const int *types, const double *values, double *phi, const double *u,
                         const double *v, const double *w, int nsteps);
/* lsf_mass (wrapper_func.cpp:450-511)                         [golden] */
double orc_lsf_mass(int ndim, int nx, int ny, int nz, double dv, double ls_width, double ratio,
                    int positive_ref, const double *phi);
/* `count` Euler steps with mpls (source.cpp:109-146)          [golden] */
void orc_mpls_euler(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                    double ls_width, double ratio, int positive_ref, double mass0, const int *types,
                    const double *values, double *phi, int count, double *delta_out);
/* find_sign (lsm.cpp:33-43)                                   [golden] */
void orc_find_sign(long n, const double *phi, double ls_width, double *sign);
/* `nsteps` tvd_rk3 steps with lsf_redistance_lambda (source.cpp:148-240,
 * godunov_gradient.cpp:7-201)                                 [golden] */
void orc_redistance_rk3(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                        double rdt, const double *sign, const double *delta, const int *types,
                        const double *values, double *phi, int nsteps);

/* stabilized_upon_gradient (godunov_gradient.cpp:159-181): phi /= the
 * largest Godunov gradient (godunov_gradient.cpp:7-152, stored sign) over
 * the padded array; grad receives the gradient; returns the maximum.
 *                                                             [golden] */
double orc_stabilize_gradient(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                              const double *sign, double *phi, double *grad);

#ifdef __cplusplus
}
#endif
#endif
