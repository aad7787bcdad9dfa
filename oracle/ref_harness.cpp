// ref_harness.cpp -- C entry points over the compiled SciHPC reference.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's own scihpc/*.cpp (read in place, never copied) into
// oracle/_ref/libscihpc_ref.so.  It lets tests/golden/make_golden.py write
// golden vectors and lets tests compare the CPU restatement in
// oracle/oracle.c against the reference function by function.  Every entry
// point below only constructs reference objects through their public
// constructors and calls their public methods.
#include <cstring>
#include <string>

#include "bc_factory.h"
#include "godunov_gradient.h"
#include "multigrid.h"
#include "projection_method.h"
#include "runge_kutta.h"
#include "lsm.h"
#include "source.h"
#include "vtkWriter.h"
#include "scalar_data.h"
#include "shared_solvers.h"
#include "structured_grid.h"
#include "vector_data.h"
#include "weno_solver.h"
#include "wrapper.h"
#include "wrapper_func.h"

namespace {

const std::string kTypes[5] = {NEUMANN, DIRICHLET, PERIODIC, NO_SLIP, SLIP};

structured_grid *make_grid(int ndim, int nx, int ny, int nz, const double *ext) {
  if (ndim == 1) return new structured_grid(axis{ext[0], ext[1], nx});
  if (ndim == 2)
    return new structured_grid(axis{ext[0], ext[1], nx}, axis{ext[2], ext[3], ny});
  return new structured_grid(axis{ext[0], ext[1], nx}, axis{ext[2], ext[3], ny},
                             axis{ext[4], ext[5], nz});
}

scalar_data *make_scalar(int ndim, int nx, int ny, int nz) {
  if (ndim == 1) return new scalar_data(nx);
  if (ndim == 2) return new scalar_data(nx, ny);
  return new scalar_data(nx, ny, nz);
}

void to_ref(const scalar_data *f, DataType ***dst, const double *src) {
  long c = 0;
  for (int i = 0; i < f->Nx; ++i)
    for (int j = 0; j < f->Ny; ++j)
      for (int k = 0; k < f->Nz; ++k) dst[i][j][k] = src[c++];
}

void from_ref(const scalar_data *f, DataType ***src, double *dst) {
  long c = 0;
  for (int i = 0; i < f->Nx; ++i)
    for (int j = 0; j < f->Ny; ++j)
      for (int k = 0; k < f->Nz; ++k) dst[c++] = src[i][j][k];
}

bc_info info(const int *types, const double *values, int q) {
  bc_info b;
  b.type = kTypes[types[q]];
  b.value = values[q];
  return b;
}

wrapper *make_wrapper(bool scalar, structured_grid *geo, int ndim,
                      const int *types, const double *values) {
  if (ndim == 1) return new wrapper(scalar, geo, info(types, values, 0), info(types, values, 1));
  if (ndim == 2)
    return new wrapper(scalar, geo, info(types, values, 0), info(types, values, 1),
                       info(types, values, 2), info(types, values, 3));
  return new wrapper(scalar, geo, info(types, values, 0), info(types, values, 1),
                     info(types, values, 2), info(types, values, 3), info(types, values, 4),
                     info(types, values, 5));
}

struct RefMG {
  structured_grid *geo;
  scalar_data *var;
  multigrid *mg;
};

// convection (source.cpp:8-40) with the WENO derivative call that the
// reference keeps commented out at source.cpp:11.
void convection_weno(wrapper *f, wrapper *vel, DataType ***s,
                     void (*flux)(scalar_data *, vector_data *)) {
  flux(f->scalar, vel->vector);
  f->solvers->weno->weno5_find_derivatives(f->scalar, vel->vector);
  for (int i = 0; i < f->scalar->Nx; ++i)
    for (int j = 0; j < f->scalar->Ny; ++j)
      for (int k = 0; k < f->scalar->Nz; ++k)
        s[i][j][k] = -vel->vector->x.data[i][j][k] * f->scalar->fx[i][j][k];
  if (f->scalar->ndim > 1)
    for (int i = 0; i < f->scalar->Nx; ++i)
      for (int j = 0; j < f->scalar->Ny; ++j)
        for (int k = 0; k < f->scalar->Nz; ++k)
          s[i][j][k] -= vel->vector->y.data[i][j][k] * f->scalar->fy[i][j][k];
  if (f->scalar->ndim > 2)
    for (int i = 0; i < f->scalar->Nx; ++i)
      for (int j = 0; j < f->scalar->Ny; ++j)
        for (int k = 0; k < f->scalar->Nz; ++k)
          s[i][j][k] -= vel->vector->z.data[i][j][k] * f->scalar->fz[i][j][k];
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- multigrid
void *ref_mg_create(int ndim, int nx, int ny, int nz, const double *ext, int bc) {
  auto *r = new RefMG;
  r->geo = make_grid(ndim, nx, ny, nz, ext);
  r->var = make_scalar(ndim, nx, ny, nz);
  r->mg = new multigrid(r->var, r->geo);
  for (unsigned i = 0; i < r->mg->level_num; ++i) {
    if (bc == 0) {
      r->mg->at[i]->init_NeumannBC();
    } else {
      r->mg->at[i]->init_DirichletBC();
      r->mg->at[i]->no_compatibility = true;
    }
    r->mg->at[i]->rhs.setZero();
    r->mg->at[i]->sol.setZero();
    r->mg->at[i]->res.setZero();
  }
  return r;
}

int ref_mg_nlevels(void *h) { return (int)static_cast<RefMG *>(h)->mg->level_num; }

void ref_mg_level_shape(void *h, int level, int *shape) {
  auto *b = static_cast<RefMG *>(h)->mg->at[level];
  shape[0] = b->nx;
  shape[1] = b->ny;
  shape[2] = b->nz;
  shape[3] = b->degree;
  shape[4] = b->n;
}

void ref_mg_set(void *h, int level, int which, const double *src) {
  auto *b = static_cast<RefMG *>(h)->mg->at[level];
  VectorX &v = which == 0 ? b->rhs : which == 1 ? b->sol : b->res;
  for (int i = 0; i < b->n; ++i) v[i] = src[i];
}

void ref_mg_get(void *h, int level, int which, double *dst) {
  auto *b = static_cast<RefMG *>(h)->mg->at[level];
  VectorX &v = which == 0 ? b->rhs : which == 1 ? b->sol : which == 2 ? b->res : b->buffer;
  for (int i = 0; i < b->n; ++i) dst[i] = v[i];
}

void ref_mg_set_nocompat(void *h, int level, int flag) {
  static_cast<RefMG *>(h)->mg->at[level]->no_compatibility = flag != 0;
}

void ref_mg_apply(void *h, int level, const double *x, double *y) {
  auto *b = static_cast<RefMG *>(h)->mg->at[level];
  VectorX xv(b->n);
  for (int i = 0; i < b->n; ++i) xv[i] = x[i];
  b->Ax(xv);
  for (int i = 0; i < b->n; ++i) y[i] = b->buffer[i];
}

double ref_mg_residual(void *h, int level) { return static_cast<RefMG *>(h)->mg->at[level]->residual(); }
void ref_mg_relax(void *h, int level, int iter) { static_cast<RefMG *>(h)->mg->at[level]->relax(iter); }
void ref_mg_restriction(void *h, int level) {
  auto *m = static_cast<RefMG *>(h)->mg;
  m->at[level]->restriction(m->at[level + 1]);
}
void ref_mg_prolongation(void *h, int level) {
  auto *m = static_cast<RefMG *>(h)->mg;
  m->at[level]->prolongation(m->at[level - 1]);
}
void ref_mg_solve(void *h) {
  auto *m = static_cast<RefMG *>(h)->mg;
  m->at[m->level_num - 1]->solve();
}
void ref_mg_vcycle(void *h) { static_cast<RefMG *>(h)->mg->v_cycle(); }
void ref_mg_fullcycle(void *h) { static_cast<RefMG *>(h)->mg->full_cycle(); }

// ------------------------------------------------------------------ BCs
void ref_bc_apply(int ndim, int nx, int ny, int nz, const double *ext, const int *types,
                  const double *values, int stag, double *f) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  scalar_data *s = make_scalar(ndim, nx, ny, nz);
  bc_factory *bf;
  if (ndim == 1)
    bf = new bc_factory(geo, info(types, values, 0), info(types, values, 1));
  else if (ndim == 2)
    bf = new bc_factory(geo, info(types, values, 0), info(types, values, 1), info(types, values, 2),
                        info(types, values, 3));
  else
    bf = new bc_factory(geo, info(types, values, 0), info(types, values, 1), info(types, values, 2),
                        info(types, values, 3), info(types, values, 4), info(types, values, 5));
  to_ref(s, s->data, f);
  if (stag < 0)
    bf->apply_node_centered(s);
  else if (stag == 0)
    bf->apply_node_staggered_x(s);
  else if (stag == 1)
    bf->apply_node_staggered_y(s);
  else
    bf->apply_node_staggered_z(s);
  from_ref(s, s->data, f);
}

// ----------------------------------------------------------------- WENO
void ref_weno_derivatives(int ndim, int nx, int ny, int nz, const double *ext,
                          const double *flux, const double *u, const double *v,
                          const double *w, double *fx, double *fy, double *fz) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  scalar_data *f = make_scalar(ndim, nx, ny, nz);
  vector_data *vel = ndim == 1   ? new vector_data(nx)
                     : ndim == 2 ? new vector_data(nx, ny)
                                 : new vector_data(nx, ny, nz);
  to_ref(f, f->flux, flux);
  to_ref(f, vel->x.data, u);
  to_ref(f, vel->y.data, v);
  to_ref(f, vel->z.data, w);
  weno_solver ws(f, geo);
  ws.weno5_find_derivatives(f, vel);
  from_ref(f, f->fx, fx);
  from_ref(f, f->fy, fy);
  from_ref(f, f->fz, fz);
}

// ------------------------------------------------- TVD-RK3 + WENO advection
void ref_tvd_rk3_weno(int ndim, int nx, int ny, int nz, const double *ext, double dt,
                      const int *types, const double *values, double *phi, const double *u,
                      const double *v, const double *w, int nsteps) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  wrapper *f = make_wrapper(true, geo, ndim, types, values);
  int vt[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *vel = make_wrapper(false, geo, ndim, vt, vv);
  auto *param = new problem_parameters();
  param->dt = dt;
  auto *solvers = SharedSolvers_alloc(f->scalar, geo);
  auto *dummy = dummy_data_alloc(f->scalar);
  f->link_params(param);
  f->link_solvers(solvers);
  f->link_dummy(dummy);
  vel->link_params(param);
  vel->link_solvers(solvers);
  vel->link_dummy(dummy);
  to_ref(f->scalar, f->scalar->data, phi);
  to_ref(f->scalar, vel->vector->x.data, u);
  to_ref(f->scalar, vel->vector->y.data, v);
  to_ref(f->scalar, vel->vector->z.data, w);
  runge_kutta rk(f->scalar->Nx, f->scalar->Ny, f->scalar->Nz);
  for (int s = 0; s < nsteps; ++s) rk.tvd_rk3(f, vel, &identity_flux, &convection_weno);
  from_ref(f->scalar, f->scalar->data, phi);
}

// -------------------------------------------------------------- density
void ref_density(int ndim, int nx, int ny, int nz, const double *ext, const double *phi,
                 double ls_width, double ratio, int positive_ref, double *rho) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  int vt[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *lsf = make_wrapper(true, geo, ndim, vt, vv);
  auto *param = new problem_parameters();
  param->ls_width = ls_width;
  param->density_ratio = ratio;
  param->positive_ref = positive_ref != 0;
  lsf->link_params(param);
  lsf->link_dummy(dummy_data_alloc(lsf->scalar));
  to_ref(lsf->scalar, lsf->scalar->data, phi);
  find_density(lsf);
  from_ref(lsf->scalar, lsf->dummy->density, rho);
}

// ------------------------------------------------- projection (SOR path)
// init_ppe_proj_coefficients + projection(): RHS, SOR sweeps, velocity
// update, velocity BCs.  Outputs C* and RHS arrays, the pressure and the
// face velocities (all padded).
void ref_projection(int ndim, int nx, int ny, int nz, const double *ext, double dt,
                    int max_iter, double tol, double tol2, const double *rho, double *p,
                    double *u, double *v, double *w, double *coef /* 8 padded arrays */) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  int nt[6] = {0, 0, 0, 0, 0, 0};
  int st[6] = {3, 3, 3, 3, 3, 3};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *pressure = make_wrapper(true, geo, ndim, nt, vv);
  wrapper *lsf = make_wrapper(true, geo, ndim, nt, vv);
  wrapper *vel = make_wrapper(false, geo, ndim, st, vv);
  auto *param = new problem_parameters();
  param->dt = dt;
  param->ppe_max_iter = max_iter;
  param->ppe_tol = tol;
  param->ppe_tol2 = tol2;
  auto *dummy = dummy_data_alloc(lsf->scalar);
  for (wrapper *x : {pressure, lsf, vel}) {
    x->link_params(param);
    x->link_dummy(dummy);
  }
  scalar_data *s = lsf->scalar;
  to_ref(s, dummy->density, rho);
  to_ref(s, pressure->scalar->data, p);
  to_ref(s, vel->vector->x.data, u);
  to_ref(s, vel->vector->y.data, v);
  to_ref(s, vel->vector->z.data, w);
  projection_method pm(s);
  pm.init_ppe_proj_coefficients(pressure);
  pm.projection(pressure, lsf, vel);
  from_ref(s, pressure->scalar->data, p);
  from_ref(s, vel->vector->x.data, u);
  from_ref(s, vel->vector->y.data, v);
  from_ref(s, vel->vector->z.data, w);
  long tot = (long)s->Nx * s->Ny * s->Nz;
  DataType ***arrs[8] = {pm.CC, pm.CR, pm.CL, pm.CU, pm.CD, pm.CF, pm.CB, pm.RHS};
  for (int a = 0; a < 8; ++a) from_ref(s, arrs[a], coef + a * tot);
}

// ------------------------------------------------------ FPC coefficients
void ref_fpc_coefficients(int ndim, int nx, int ny, int nz, const double *ext, double rho12,
                          const double *rho, const double *phat, double *coef) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  int nt[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *pressure = make_wrapper(true, geo, ndim, nt, vv);
  auto *param = new problem_parameters();
  param->density_ratio = rho12;
  auto *dummy = dummy_data_alloc(pressure->scalar);
  pressure->link_params(param);
  pressure->link_dummy(dummy);
  scalar_data *s = pressure->scalar;
  to_ref(s, dummy->density, rho);
  to_ref(s, s->flux, phat);
  projection_method pm(s);
  pm.init_ppe_fpc_coefficients(pressure);
  long tot = (long)s->Nx * s->Ny * s->Nz;
  DataType ***arrs[8] = {pm.CC, pm.CR, pm.CL, pm.CU, pm.CD, pm.CF, pm.CB, pm.RHS};
  for (int a = 0; a < 8; ++a) from_ref(s, arrs[a], coef + a * tot);
}

// -------------------------------------------------------------- FPC step
// pressure_guess_solution + init_ppe_fpc_coefficients +
// fast_pressure_correction on a fresh pressure wrapper (Neumann BCs, its
// multigrid starting from p = 0) and face velocities with the given BC
// types.  Outputs the pressure, p_hat (the pressure's flux array), the
// corrected velocities (all padded) and the multigrid cycles.
int ref_fpc_step(int ndim, int nx, int ny, int nz, const double *ext, double dt, double dt_old,
                 double dt_old2, double rho12, double ppe_tol, const int *vel_types, const double *rho,
                 const double *p_old, const double *p_old2, double *u, double *v, double *w, double *p,
                 double *phat) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  int nt[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *pressure = make_wrapper(true, geo, ndim, nt, vv);
  wrapper *lsf = make_wrapper(true, geo, ndim, nt, vv);
  wrapper *vel = make_wrapper(false, geo, ndim, vel_types, vv);
  auto *param = new problem_parameters();
  param->dt = dt;
  param->dt_old = dt_old;
  param->dt_old2 = dt_old2;
  param->density_ratio = rho12;
  param->ppe_tol = ppe_tol;
  auto *dummy = dummy_data_alloc(lsf->scalar);
  auto *solvers = SharedSolvers_alloc(pressure->scalar, geo);
  shared_solvers_mg_init_Neumann(solvers);
  for (wrapper *x : {pressure, lsf, vel}) {
    x->link_params(param);
    x->link_dummy(dummy);
  }
  pressure->link_solvers(solvers);
  scalar_data *s = pressure->scalar;
  to_ref(s, dummy->density, rho);
  to_ref(s, s->old, p_old);
  to_ref(s, s->old2, p_old2);
  to_ref(s, vel->vector->x.data, u);
  to_ref(s, vel->vector->y.data, v);
  to_ref(s, vel->vector->z.data, w);
  projection_method pm(s);
  pm.pressure_guess_solution(pressure);
  pm.init_ppe_fpc_coefficients(pressure);
  pm.fast_pressure_correction(pressure, lsf, vel);
  from_ref(s, s->data, p);
  from_ref(s, s->flux, phat);
  from_ref(s, vel->vector->x.data, u);
  from_ref(s, vel->vector->y.data, v);
  from_ref(s, vel->vector->z.data, w);
  return 0;
}

// vtkWriter(grid, prefix).create(id), add_scalar per padded scalar field,
// add_vector per padded (u, v, w) triple (vectors[3q+c]), close().
int ref_vtk_write(const char *prefix, unsigned id, int ndim, int nx, int ny, int nz,
                  const double *ext, int nscalar, const char *const *snames,
                  const double *const *scalars, int nvector, const char *const *vnames,
                  const double *const *vectors) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  vtkWriter vtk(geo, prefix);
  vtk.create(id);
  for (int q = 0; q < nscalar; ++q) {
    scalar_data *f = make_scalar(ndim, nx, ny, nz);
    to_ref(f, f->data, scalars[q]);
    vtk.add_scalar(f, snames[q]);
  }
  for (int q = 0; q < nvector; ++q) {
    vector_data *v = ndim == 1 ? new vector_data(nx)
                     : ndim == 2 ? new vector_data(nx, ny)
                                 : new vector_data(nx, ny, nz);
    to_ref(&v->x, v->x.data, vectors[3 * q]);
    to_ref(&v->x, v->y.data, vectors[3 * q + 1]);
    to_ref(&v->x, v->z.data, vectors[3 * q + 2]);
    vtk.add_vector(v, vnames[q]);
  }
  vtk.close();
  return 0;
}

// projection_method::find_source_sec + find_intermediate_velocity on fresh
// wrappers: vel and nvel with the given BC types (both), lsf with phi,
// parameters prm8 = {dt, Re, Fr, We, ls_width, density_ratio,
// viscosity_ratio, positive_ref}.  src_* are u_src/v_src/w_src on entry
// (the previous step's sources) and on exit; nvel_* are the node
// velocities on entry and the last face transfer on exit.
int ref_momentum_step(int ndim, int nx, int ny, int nz, const double *ext, const double *prm8,
                      const int *vel_types, const double *phi, double *u, double *v, double *w,
                      double *nu, double *nv, double *nw, double *su, double *sv, double *sw,
                      double *rho_out) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  int nt[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *lsf = make_wrapper(true, geo, ndim, nt, vv);
  wrapper *vel = make_wrapper(false, geo, ndim, vel_types, vv);
  wrapper *nvel = make_wrapper(false, geo, ndim, vel_types, vv);
  auto *param = new problem_parameters();
  param->dt = prm8[0];
  param->Reynolds_number = prm8[1];
  param->Froude_number = prm8[2];
  param->Weber_number = prm8[3];
  param->ls_width = prm8[4];
  param->density_ratio = prm8[5];
  param->viscosity_ratio = prm8[6];
  param->positive_ref = prm8[7] != 0.0;
  auto *dummy = dummy_data_alloc(lsf->scalar);
  auto *solvers = SharedSolvers_alloc(lsf->scalar, geo);
  for (wrapper *x : {lsf, vel, nvel}) {
    x->link_params(param);
    x->link_dummy(dummy);
    x->link_solvers(solvers);
  }
  scalar_data *s = lsf->scalar;
  to_ref(s, s->data, phi);
  to_ref(s, vel->vector->x.data, u);
  to_ref(s, vel->vector->y.data, v);
  to_ref(s, vel->vector->z.data, w);
  to_ref(s, nvel->vector->x.data, nu);
  to_ref(s, nvel->vector->y.data, nv);
  to_ref(s, nvel->vector->z.data, nw);
  projection_method pm(s);
  to_ref(s, pm.u_src, su);
  to_ref(s, pm.v_src, sv);
  to_ref(s, pm.w_src, sw);
  pm.find_source_sec(vel, nvel, lsf);
  pm.find_intermediate_velocity(vel);
  from_ref(s, vel->vector->x.data, u);
  from_ref(s, vel->vector->y.data, v);
  from_ref(s, vel->vector->z.data, w);
  from_ref(s, nvel->vector->x.data, nu);
  from_ref(s, nvel->vector->y.data, nv);
  from_ref(s, nvel->vector->z.data, nw);
  from_ref(s, pm.u_src, su);
  from_ref(s, pm.v_src, sv);
  from_ref(s, pm.w_src, sw);
  from_ref(s, dummy->density, rho_out);
  return 0;
}

// node_from_face + nvel->apply_nvel_bc (the end of ab_solve_sec)
int ref_node_velocity(int ndim, int nx, int ny, int nz, const double *ext, const int *vel_types,
                      const double *u, const double *v, const double *w, double *nu, double *nv,
                      double *nw) {
  structured_grid *geo = make_grid(ndim, nx, ny, nz, ext);
  double vv[6] = {0, 0, 0, 0, 0, 0};
  wrapper *vel = make_wrapper(false, geo, ndim, vel_types, vv);
  wrapper *nvel = make_wrapper(false, geo, ndim, vel_types, vv);
  scalar_data *s = &vel->vector->x;
  to_ref(s, vel->vector->x.data, u);
  to_ref(s, vel->vector->y.data, v);
  to_ref(s, vel->vector->z.data, w);
  to_ref(s, nvel->vector->x.data, nu);
  to_ref(s, nvel->vector->y.data, nv);
  to_ref(s, nvel->vector->z.data, nw);
  node_from_face(vel, nvel);
  nvel->apply_nvel_bc();
  from_ref(s, nvel->vector->x.data, nu);
  from_ref(s, nvel->vector->y.data, nv);
  from_ref(s, nvel->vector->z.data, nw);
  return 0;
}

// ------------------------------------------------- interface maintenance
namespace {
struct LsSetup {
  structured_grid *geo;
  wrapper *f, *vel;
  problem_parameters *param;
  dummy_data *dummy;
};
LsSetup ls_setup(int ndim, int nx, int ny, int nz, const double *ext, const int *types, const double *values,
                 const double *phi, const double *u, const double *v, const double *w) {
  LsSetup L;
  L.geo = make_grid(ndim, nx, ny, nz, ext);
  L.f = make_wrapper(true, L.geo, ndim, types, values);
  int vt[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  L.vel = make_wrapper(false, L.geo, ndim, vt, vv);
  L.param = new problem_parameters();
  auto *solvers = SharedSolvers_alloc(L.f->scalar, L.geo);
  L.dummy = dummy_data_alloc(L.f->scalar);
  for (wrapper *x : {L.f, L.vel}) {
    x->link_params(L.param);
    x->link_solvers(solvers);
    x->link_dummy(L.dummy);
  }
  scalar_data *s = L.f->scalar;
  to_ref(s, s->data, phi);
  if (u) to_ref(s, L.vel->vector->x.data, u);
  if (v) to_ref(s, L.vel->vector->y.data, v);
  if (w) to_ref(s, L.vel->vector->z.data, w);
  return L;
}
}  // namespace

// uccd_solver::find_derivatives(f, vel) (upwind = 1) or
// ccd_solver::find_derivatives(f) on the flux array
void ref_compact_derivatives(int ndim, int nx, int ny, int nz, const double *ext, const double *flux,
                             const double *u, const double *v, const double *w, int upwind, double *fx,
                             double *fy, double *fz, double *fxx, double *fyy, double *fzz) {
  int t[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, t, vv, flux, u, v, w);
  scalar_data *s = L.f->scalar;
  to_ref(s, s->flux, flux);
  if (upwind) L.f->solvers->uccd->find_derivatives(s, L.vel->vector);
  else L.f->solvers->ccd->find_derivatives(s);
  from_ref(s, s->fx, fx);
  from_ref(s, s->fxx, fxx);
  if (ndim > 1) {
    from_ref(s, s->fy, fy);
    from_ref(s, s->fyy, fyy);
  }
  if (ndim > 2) {
    from_ref(s, s->fz, fz);
    from_ref(s, s->fzz, fzz);
  }
}

// runge_kutta::tvd_rk3(phi, nvel, identity_flux, convection) (UCCD)
void ref_tvd_rk3_uccd(int ndim, int nx, int ny, int nz, const double *ext, double dt, const int *types,
                      const double *values, double *phi, const double *u, const double *v, const double *w,
                      int nsteps) {
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, types, values, phi, u, v, w);
  L.param->dt = dt;
  runge_kutta rk(L.f->scalar->Nx, L.f->scalar->Ny, L.f->scalar->Nz);
  for (int s = 0; s < nsteps; ++s) rk.tvd_rk3(L.f, L.vel, &identity_flux, &convection);
  from_ref(L.f->scalar, L.f->scalar->data, phi);
}

double ref_lsf_mass(int ndim, int nx, int ny, int nz, const double *ext, double ls_width, double ratio,
                    int positive_ref, const double *phi) {
  int t[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, t, vv, phi, nullptr, nullptr, nullptr);
  L.param->ls_width = ls_width;
  L.param->density_ratio = ratio;
  L.param->positive_ref = positive_ref != 0;
  return lsf_mass(L.f);
}

// `count` runge_kutta::euler(phi, nvel, identity_flux, mpls) steps
void ref_mpls_euler(int ndim, int nx, int ny, int nz, const double *ext, double dt, double ls_width,
                    double ratio, int positive_ref, double mass0, const int *types, const double *values,
                    double *phi, int count, double *delta_out) {
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, types, values, phi, nullptr, nullptr, nullptr);
  L.param->dt = dt;
  L.param->ls_width = ls_width;
  L.param->density_ratio = ratio;
  L.param->positive_ref = positive_ref != 0;
  L.param->lsf_mass0 = mass0;
  runge_kutta rk(L.f->scalar->Nx, L.f->scalar->Ny, L.f->scalar->Nz);
  for (int s = 0; s < count; ++s) rk.euler(L.f, L.vel, &identity_flux, &mpls);
  from_ref(L.f->scalar, L.f->scalar->data, phi);
  if (delta_out) from_ref(L.f->scalar, L.dummy->delta, delta_out);
}

// find_sign (wrapper_func.cpp:31-40)
void ref_find_sign(int ndim, int nx, int ny, int nz, const double *ext, double ls_width, const double *phi,
                   double *sign) {
  int t[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, t, vv, phi, nullptr, nullptr, nullptr);
  L.param->ls_width = ls_width;
  find_sign(L.f);
  from_ref(L.f->scalar, L.dummy->sign, sign);
}

// `nsteps` runge_kutta::tvd_rk3(phi, nvel, identity_flux,
// lsf_redistance_lambda) with dummy->sign and dummy->delta given
void ref_redistance_rk3(int ndim, int nx, int ny, int nz, const double *ext, double dt, double rdt,
                        const double *sign, const double *delta, const int *types, const double *values,
                        double *phi, int nsteps) {
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, types, values, phi, nullptr, nullptr, nullptr);
  L.param->dt = dt;
  L.param->rdt = rdt;
  to_ref(L.f->scalar, L.dummy->sign, sign);
  to_ref(L.f->scalar, L.dummy->delta, delta);
  runge_kutta rk(L.f->scalar->Nx, L.f->scalar->Ny, L.f->scalar->Nz);
  for (int s = 0; s < nsteps; ++s) rk.tvd_rk3(L.f, L.vel, &identity_flux, &lsf_redistance_lambda);
  from_ref(L.f->scalar, L.f->scalar->data, phi);
}

// stabilized_upon_gradient(phi) (godunov_gradient.cpp:159-181) with
// dummy->sign given; phi scaled in place, grad = dummy->grad
void ref_stabilize_gradient(int ndim, int nx, int ny, int nz, const double *ext, const double *sign, double *phi,
                            double *grad) {
  int t[6] = {0, 0, 0, 0, 0, 0};
  double vv[6] = {0, 0, 0, 0, 0, 0};
  LsSetup L = ls_setup(ndim, nx, ny, nz, ext, t, vv, phi, nullptr, nullptr, nullptr);
  to_ref(L.f->scalar, L.dummy->sign, sign);
  stabilized_upon_gradient(L.f);
  from_ref(L.f->scalar, L.f->scalar->data, phi);
  from_ref(L.f->scalar, L.dummy->grad, grad);
}

}  // extern "C"
