// scihpc_facade.hpp -- SciHPC's solver classes over the MI355X C ABI.
//
// Header-only C++17.  The classes keep the public interface that SciHPC's
// drivers and projection step use (paths relative to the SciHPC tree):
//   structured_grid   structured_grid.h:10-30   (axis, dx, xc, ...)
//   scalar_data       scalar_data.h:10-32       (ghc = 3, data[i][j][k], index_mapping)
//   vector_data       vector_data.h             (x, y, z components)
//   multigrid_base    multigrid_base.h:17-55    (rhs/sol/res, of, residual, relax, ...)
//   multigrid         multigrid.h:13-25         (level_num, at[], v_cycle, full_cycle)
//   weno_solver       weno_solver.h:17-48       (weno5_find_derivatives)
//   runge_kutta       runge_kutta.h:12-23       (tvd_rk3 of the WENO5 level-set convection)
//   projection_method projection_method.h:14-51 (fast_pressure_correction, projection)
// so a driver written against SciHPC compiles against this header with
// `using namespace scihpc_gpu;`.  Every operation runs on the GPU through
// include/scihpc_hip.h; host vectors are copied to the device before a call
// that reads them and back after a call that writes them.  Errors become
// scihpc_gpu::error exceptions carrying the ABI status and message.
//
// Differences a caller sees:
//  * the multigrid boundary operator is per hierarchy: init_NeumannBC() /
//    init_DirichletBC() on any level selects it for all levels (every SciHPC
//    driver sets the same one on every level);
//  * runge_kutta::tvd_rk3 advances the level set with the WENO5 convection
//    right-hand side and the identity flux (the combination SciHPC's
//    level-set step uses, source.cpp:8-40, flux.cpp:7-30) instead of taking
//    arbitrary host callbacks, which cannot run on the device;
//  * multigrid::solve() is added: the variable-density A(rho) solve to a
//    relative tolerance (FP64 or mixed precision, optional learned guess).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "scihpc_hip.h"

namespace scihpc_gpu {

using DataType = double;

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != SCIHPC_OK && rc != SCIHPC_NOT_CONVERGED)
    throw error(rc, std::string("scihpc: ") + scihpc_last_error());
}

struct axis {
  DataType start, end;
  int n;
};

struct indices {
  int i, j, k;
};

// structured_grid.cpp:43-133: n cells per axis, nodes x[0..n], centres xc
class structured_grid {
 public:
  explicit structured_grid(axis xa) : structured_grid(1, xa, axis{0, 0, 1}, axis{0, 0, 1}) {}
  structured_grid(axis xa, axis ya) : structured_grid(2, xa, ya, axis{0, 0, 1}) {}
  structured_grid(axis xa, axis ya, axis za) : structured_grid(3, xa, ya, za) {}

  DataType *x, *xc, *y, *yc, *z, *zc;
  DataType dx = 0, dy = 0, dz = 0, h = 0, dv = 0;
  axis x_axis, y_axis, z_axis;
  int ndim;

 private:
  std::vector<DataType> store_[6];
  structured_grid(int nd, axis xa, axis ya, axis za) : x_axis(xa), y_axis(ya), z_axis(za), ndim(nd) {
    auto fill = [&](const axis &a, int s, DataType &d, bool used) {
      d = used ? (a.end - a.start) / static_cast<DataType>(a.n) : 0.0;
      store_[s].resize(a.n + 1);
      store_[s + 1].resize(a.n);
      for (int i = 0; i <= a.n; ++i) store_[s][i] = a.start + static_cast<DataType>(i * d);
      for (int i = 0; i < a.n; ++i) store_[s + 1][i] = store_[s][i] + static_cast<DataType>(0.5 * d);
    };
    fill(xa, 0, dx, true);
    fill(ya, 2, dy, nd >= 2);
    fill(za, 4, dz, nd >= 3);
    x = store_[0].data();
    xc = store_[1].data();
    y = store_[2].data();
    yc = store_[3].data();
    z = store_[4].data();
    zc = store_[5].data();
    h = nd == 1 ? dx : nd == 2 ? std::fmin(dx, dy) : std::fmin(std::fmin(dx, dy), dz);
    dv = nd == 1 ? dx : nd == 2 ? dx * dy : dx * dy * dz;
  }
};

// scalar_data.h:10-32: padded fields data[I][J][K] (ghost 3 on axes with
// more than one cell).  Storage is contiguous, k fastest -- the padded
// layout of the C ABI -- behind the same triple-pointer access.
class scalar_data {
 public:
  int ghc = 3;
  int Nx, Ny, Nz, nx, ny, nz, ndim;
  DataType ***data, ***flux, ***old, ***old2, ***fx, ***fy, ***fz;

  explicit scalar_data(int _nx) : scalar_data(_nx, 1, 1, 1) {}
  scalar_data(int _nx, int _ny) : scalar_data(_nx, _ny, 1, 2) {}
  scalar_data(int _nx, int _ny, int _nz) : scalar_data(_nx, _ny, _nz, 3) {}

  // scalar_data.cpp:95-123 (1-based interior indices)
  indices index_mapping(int i, int j, int k) const {
    return indices{i - 1 + (nx > 1 ? ghc : 0), j - 1 + (ny > 1 ? ghc : 0), k - 1 + (nz > 1 ? ghc : 0)};
  }
  long size() const { return (long)Nx * Ny * Nz; }
  // contiguous base of a field
  static DataType *base(DataType ***f) { return &f[0][0][0]; }

 private:
  struct field {
    std::vector<DataType> v;
    std::vector<DataType *> rows;
    std::vector<DataType **> planes;
  };
  field f_[7];
  DataType ***make(field &f) {
    f.v.assign((size_t)Nx * Ny * Nz, 0.0);
    f.rows.resize((size_t)Nx * Ny);
    f.planes.resize(Nx);
    for (int i = 0; i < Nx; ++i) {
      for (int j = 0; j < Ny; ++j) f.rows[(size_t)i * Ny + j] = f.v.data() + ((size_t)i * Ny + j) * Nz;
      f.planes[i] = f.rows.data() + (size_t)i * Ny;
    }
    return f.planes.data();
  }
  scalar_data(int _nx, int _ny, int _nz, int nd) : nx(_nx), ny(_ny), nz(_nz), ndim(nd) {
    Nx = nx > 1 ? nx + 2 * ghc : 1;
    Ny = ny > 1 ? ny + 2 * ghc : 1;
    Nz = nz > 1 ? nz + 2 * ghc : 1;
    data = make(f_[0]);
    flux = make(f_[1]);
    old = make(f_[2]);
    old2 = make(f_[3]);
    fx = make(f_[4]);
    fy = make(f_[5]);
    fz = make(f_[6]);
  }

 public:
  // scalar_data::store (scalar_data.cpp): old2 <- old <- data
  void store() {
    std::copy(f_[2].v.begin(), f_[2].v.end(), f_[3].v.begin());
    std::copy(f_[0].v.begin(), f_[0].v.end(), f_[2].v.begin());
  }
};

class vector_data {
 public:
  std::unique_ptr<scalar_data> x, y, z;
  explicit vector_data(int _nx) : x(new scalar_data(_nx)) {}
  vector_data(int _nx, int _ny) : x(new scalar_data(_nx, _ny)), y(new scalar_data(_nx, _ny)) {}
  vector_data(int _nx, int _ny, int _nz)
      : x(new scalar_data(_nx, _ny, _nz)), y(new scalar_data(_nx, _ny, _nz)), z(new scalar_data(_nx, _ny, _nz)) {}
};

// Eigen::VectorX stand-in for the multigrid's host vectors
class VectorX {
 public:
  explicit VectorX(long n = 0) : v_(n, 0.0) {}
  DataType &operator[](long i) { return v_[i]; }
  DataType operator[](long i) const { return v_[i]; }
  DataType &operator()(long i) { return v_[i]; }
  long size() const { return (long)v_.size(); }
  void setZero() { std::fill(v_.begin(), v_.end(), 0.0); }
  DataType *data() { return v_.data(); }
  const DataType *data() const { return v_.data(); }

 private:
  std::vector<DataType> v_;
};

class multigrid;

// multigrid_base.h:17-55: one level of the hierarchy
class multigrid_base {
 public:
  VectorX rhs, sol, res;
  int nx, ny, nz, degree, ndim, n;
  DataType dx, dy, dz;
  bool no_compatibility = false;
  int base_step = 10;

  int of(int i, int j, int k) const { return i + nx * (j + ny * k); }  // multigrid_base.cpp:8-14
  int of(int i, int j) const { return i + nx * j; }

  void init_NeumannBC();
  void init_DirichletBC();
  DataType residual();                                  // multigrid_base.cpp:69-72
  void relax(int iter);                                 // multigrid_base.cpp:16-45
  void compatibility_condition(VectorX &f);             // multigrid_base.cpp:47-67
  void restriction(multigrid_base *coarse);             // multigrid_base.cpp:74-113
  void prolongation(multigrid_base *dense);             // multigrid_base.cpp:115-208
  void solve();                                         // multigrid_base.cpp:210-234
  void Ax(const VectorX &x, VectorX &out);              // multigrid_base.cpp:388-390

 private:
  friend class multigrid;
  multigrid *mg_ = nullptr;
  int level_ = 0;
  void push();
  void pull_sol();
};

// multigrid.h:13-25
class multigrid {
 public:
  multigrid(scalar_data *var, structured_grid *geo) : ndim_(var->ndim), nx_(var->nx), ny_(var->ny), nz_(var->nz),
                                                       dx_(geo->dx), dy_(geo->dy), dz_(geo->dz) {
    create(SCIHPC_MG_NEUMANN);
    int L = 0;
    check(scihpc_mg_num_levels(h_, &L));
    level_num = (unsigned)L;
    levels_.resize(L);
    ptrs_.resize(L);
    for (int l = 0; l < L; ++l) {
      auto &b = levels_[l];
      int a, c, e, deg;
      check(scihpc_mg_level_shape(h_, l, &a, &c, &e, &deg));
      b.nx = a;
      b.ny = c;
      b.nz = e;
      b.degree = deg;
      b.ndim = ndim_;
      b.n = a * c * e;
      b.dx = dx_ * deg;
      b.dy = dy_ * deg;
      b.dz = dz_ * deg;
      b.rhs = VectorX(b.n);
      b.sol = VectorX(b.n);
      b.res = VectorX(b.n);
      b.mg_ = this;
      b.level_ = l;
      ptrs_[l] = &levels_[l];
    }
    at = ptrs_.data();
  }
  ~multigrid() {
    if (h_) scihpc_mg_destroy(h_);
  }
  multigrid(const multigrid &) = delete;
  multigrid &operator=(const multigrid &) = delete;

  unsigned int level_num = 0;
  multigrid_base **at = nullptr;

  void v_cycle() {  // multigrid.cpp:43-55
    at[0]->push();
    check(scihpc_mg_vcycle_ref(h_));
    at[0]->pull_sol();
  }
  void full_cycle() {  // multigrid.cpp:57-79
    at[0]->push();
    check(scihpc_mg_fullcycle_ref(h_));
    at[0]->pull_sol();
  }

  // A(rho) from a padded density field (init_ppe_proj_coefficients,
  // projection_method.cpp:404-465)
  void set_density(const scalar_data *rho) {
    check(scihpc_mg_set_density_padded(h_, &rho->data[0][0][0]));
  }
  // solve A p = at[0]->rhs to ||r|| / ||b|| <= rel_tol (returns the ABI stats)
  scihpc_solve_stats solve(double rel_tol, int precision = SCIHPC_PREC_MIXED, int accel = 1, int max_cycles = 200,
                           int use_learned = 0) {
    at[0]->push();
    scihpc_solve_opts o{rel_tol, max_cycles, 2, 2, precision, use_learned, 1, accel};
    scihpc_solve_stats st{};
    check(scihpc_mg_solve(h_, &o, &st));
    at[0]->pull_sol();
    return st;
  }
  scihpc_mg *handle() const { return h_; }

 private:
  friend class multigrid_base;
  int ndim_, nx_, ny_, nz_;
  DataType dx_, dy_, dz_;
  int bc_ = -1;
  scihpc_mg *h_ = nullptr;
  std::vector<multigrid_base> levels_;
  std::vector<multigrid_base *> ptrs_;

  void create(int bc) {
    if (h_ && bc == bc_) return;
    if (h_) scihpc_mg_destroy(h_);
    h_ = nullptr;
    check(scihpc_mg_create(ndim_, nx_, ny_, ndim_ == 3 ? nz_ : 1, dx_, dy_, ndim_ == 3 ? dz_ : 1.0, bc, &h_));
    bc_ = bc;
    for (size_t l = 0; l < levels_.size(); ++l)
      if (levels_[l].no_compatibility) check(scihpc_mg_set_nocompat(h_, (int)l, 1));
  }
};

inline void multigrid_base::push() {
  scihpc_mg *h = mg_->h_;
  check(scihpc_mg_set_nocompat(h, level_, no_compatibility ? 1 : 0));
  check(scihpc_mg_set_rhs(h, level_, rhs.data()));
  check(scihpc_mg_set_sol(h, level_, sol.data()));
}
inline void multigrid_base::pull_sol() { check(scihpc_mg_get_sol(mg_->h_, level_, sol.data())); }
inline void multigrid_base::init_NeumannBC() { mg_->create(SCIHPC_MG_NEUMANN); }
inline void multigrid_base::init_DirichletBC() { mg_->create(SCIHPC_MG_DIRICHLET); }
inline DataType multigrid_base::residual() {
  push();
  double r = 0.0;
  check(scihpc_mg_residual(mg_->h_, level_, &r));
  check(scihpc_mg_get_res(mg_->h_, level_, res.data()));
  return r;
}
inline void multigrid_base::relax(int iter) {
  push();
  check(scihpc_mg_relax(mg_->h_, level_, iter));
  pull_sol();
}
inline void multigrid_base::compatibility_condition(VectorX &f) {
  if (no_compatibility) return;
  push();
  check(scihpc_mg_set_sol(mg_->h_, level_, f.data()));
  check(scihpc_mg_compat(mg_->h_, level_, 1));
  check(scihpc_mg_get_sol(mg_->h_, level_, f.data()));
  check(scihpc_mg_set_sol(mg_->h_, level_, sol.data()));
}
inline void multigrid_base::restriction(multigrid_base *coarse) {
  if (coarse != mg_->at[level_ + 1]) throw error(SCIHPC_BAD_ARG, "restriction target is not the next level");
  push();
  check(scihpc_mg_restriction(mg_->h_, level_));
  pull_sol();
  check(scihpc_mg_get_res(mg_->h_, level_, res.data()));
  check(scihpc_mg_get_rhs(mg_->h_, level_ + 1, coarse->rhs.data()));
  check(scihpc_mg_get_sol(mg_->h_, level_ + 1, coarse->sol.data()));
}
inline void multigrid_base::prolongation(multigrid_base *dense) {
  if (level_ == 0 || dense != mg_->at[level_ - 1]) throw error(SCIHPC_BAD_ARG, "prolongation target is not the previous level");
  push();
  dense->push();
  check(scihpc_mg_prolongation(mg_->h_, level_));
  dense->pull_sol();
}
inline void multigrid_base::solve() {
  if (level_ != (int)mg_->level_num - 1) throw error(SCIHPC_BAD_ARG, "solve() is the coarsest level's");
  push();
  check(scihpc_mg_coarse_solve(mg_->h_));
  pull_sol();
}
inline void multigrid_base::Ax(const VectorX &x, VectorX &out) {
  check(scihpc_mg_apply(mg_->h_, level_, x.data(), out.data()));
}

// boundary conditions per face (bc_factory.h:18-27): type + value
enum bc_type { NEUMANN = SCIHPC_BC_NEUMANN, DIRICHLET = SCIHPC_BC_DIRICHLET, PERIODIC = SCIHPC_BC_PERIODIC,
               NO_SLIP = SCIHPC_BC_NO_SLIP, SLIP = SCIHPC_BC_SLIP };
struct bc_info {
  bc_type type;
  DataType value = 0.0;
};

// problem_parameters (global.h:25-54): the fields the projection step reads
struct problem_parameters {
  DataType dt = 1.0, dt_old = 1.0, dt_old2 = 1.0;
  DataType density_ratio = 1.0;
  DataType ppe_tol = 1e-4;
  // momentum predictor (global.h:25-54; <= 0 disables gravity / surface tension)
  DataType Reynolds_number = 1.0, Froude_number = -1.0, Weber_number = -1.0;
  DataType ls_width = 0.0, viscosity_ratio = 1.0;
  bool positive_ref = true;
  int iter = 0;        // time step counter (the first 5 steps use projection())
  int ppe_initer = 1;  // pressure iterations per step
};

// projection_method (projection_method.h:14-51): the projection step of
// one grid on the GPU.  The multigrid is the caller's (its level-0 solution
// is the warm start of the FPC loop, as SciHPC's persistent at[0]->sol).
class projection_method {
 public:
  projection_method(scalar_data *f, structured_grid *geo, const bc_info vel_faces[6]) {
    int types[6];
    for (int q = 0; q < 6; ++q) types[q] = vel_faces[q].type;
    check(scihpc_proj_create(f->ndim, f->nx, f->ny, f->nz, geo->dx, geo->dy, f->ndim == 3 ? geo->dz : 1.0, types,
                             &h_));
  }
  ~projection_method() {
    if (h_) scihpc_proj_destroy(h_);
  }
  projection_method(const projection_method &) = delete;
  projection_method &operator=(const projection_method &) = delete;

  // pressure_guess_solution + init_ppe_fpc_coefficients +
  // fast_pressure_correction (projection_method.cpp:467-541, 659-753,
  // 803-815): pressure->old/old2 in, pressure->data and ->flux (p_hat) out,
  // vel corrected in place.  Returns the multigrid cycles.
  int fast_pressure_correction(multigrid &mg, scalar_data *pressure, const scalar_data *density, vector_data *vel,
                               const problem_parameters &prm, int max_cycles = 100) {
    scihpc_fpc_params fp{prm.dt, prm.dt_old, prm.dt_old2, prm.density_ratio, prm.ppe_tol, max_cycles};
    std::vector<double> zero;
    double *w = vel_w(vel, pressure, zero);
    int cycles = 0;
    check(scihpc_proj_fpc_step(h_, mg.handle(), &fp, base(density->data), base(pressure->old), base(pressure->old2),
                               base(vel->x->data), base(vel->y->data), w, base(pressure->data),
                               base(pressure->flux), &cycles));
    return cycles;
  }
  // find_source_sec + find_intermediate_velocity (projection_method.cpp:
  // 314-402): vel predicted in place (AB2 with the sources kept on the
  // device between calls, as the reference's u_src), nvel holds the last
  // face transfer, density (optional) receives find_density
  void find_source_sec(vector_data *vel, vector_data *nvel, const scalar_data *lsf, const problem_parameters &prm,
                       scalar_data *density = nullptr) {
    scihpc_flow_params fp{prm.dt, prm.Reynolds_number, prm.Froude_number, prm.Weber_number, prm.ls_width,
                          prm.density_ratio, prm.viscosity_ratio, prm.positive_ref ? 1 : 0};
    std::vector<double> z1, z2;
    double *w = vel_w(vel, lsf, z1), *nw = vel_w(nvel, lsf, z2);
    check(scihpc_proj_momentum_step(h_, &fp, base(lsf->data), base(vel->x->data), base(vel->y->data), w,
                                    base(nvel->x->data), base(nvel->y->data), nw,
                                    density ? base(density->data) : nullptr));
  }
  // node_from_face + nvel->apply_nvel_bc (wrapper_func.cpp:264-293)
  void node_from_face(vector_data *vel, vector_data *nvel, const scalar_data *like) {
    std::vector<double> z1, z2;
    double *w = vel_w(vel, like, z1), *nw = vel_w(nvel, like, z2);
    check(scihpc_proj_node_velocity(h_, base(vel->x->data), base(vel->y->data), w, base(nvel->x->data),
                                    base(nvel->y->data), nw));
  }
  // ab_solve_sec (projection_method.cpp:779-800): pressure history,
  // momentum predictor, projection() for the first 5 steps (here the A(rho)
  // solve) then the FPC step, node velocities.  Returns the last pressure
  // step's multigrid cycles.
  int ab_solve_sec(multigrid &mg_const, multigrid &mg_rho, vector_data *vel, vector_data *nvel,
                   scalar_data *pressure, const scalar_data *lsf, scalar_data *density,
                   const problem_parameters &prm) {
    pressure->store();
    find_source_sec(vel, nvel, lsf, prm, density);
    int cycles = 0;
    for (int cnt = 0; cnt < prm.ppe_initer; ++cnt) {
      if (prm.iter < 5) cycles = projection(mg_rho, pressure, density, vel, prm.dt).cycles;
      else cycles = fast_pressure_correction(mg_const, pressure, density, vel, prm);
    }
    node_from_face(vel, nvel, lsf);
    return cycles;
  }
  // projection() (projection_method.cpp:544-657) with the SOR replaced by the
  // A(rho) multigrid solve to rel_tol
  scihpc_solve_stats projection(multigrid &mg, scalar_data *pressure, const scalar_data *density, vector_data *vel,
                                DataType dt, double rel_tol = 1e-8, int precision = SCIHPC_PREC_MIXED) {
    scihpc_solve_opts o{rel_tol, 200, 2, 2, precision, 0, 1, 1};
    scihpc_solve_stats st{};
    std::vector<double> zero;
    double *w = vel_w(vel, pressure, zero);
    check(scihpc_proj_step(h_, mg.handle(), dt, &o, base(density->data), base(vel->x->data), base(vel->y->data), w,
                           base(pressure->data), &st));
    return st;
  }

 private:
  scihpc_proj *h_ = nullptr;
  static double *base(DataType ***f) { return &f[0][0][0]; }
  static double *vel_w(vector_data *vel, const scalar_data *like, std::vector<double> &zero) {
    if (vel->z) return base(vel->z->data);
    zero.assign(like->size(), 0.0);
    return zero.data();
  }
};

// the level-set transport of one scalar field on the GPU
class level_set_solver {
 public:
  level_set_solver(scalar_data *f, structured_grid *geo, const bc_info faces[6]) : f_(f) {
    int types[6];
    double vals[6];
    for (int q = 0; q < 6; ++q) {
      types[q] = faces[q].type;
      vals[q] = faces[q].value;
    }
    check(scihpc_ls_create(f->ndim, f->nx, f->ny, f->nz, geo->dx, geo->dy, f->ndim == 3 ? geo->dz : 1.0, types,
                           vals, &h_));
  }
  ~level_set_solver() {
    if (h_) scihpc_ls_destroy(h_);
  }
  level_set_solver(const level_set_solver &) = delete;
  level_set_solver &operator=(const level_set_solver &) = delete;
  scihpc_ls *handle() const { return h_; }
  scalar_data *field() const { return f_; }

 protected:
  scalar_data *f_;
  scihpc_ls *h_ = nullptr;
  void set_velocity(vector_data *vel) {
    std::vector<double> zero;
    const double *w = nullptr;
    if (vel->z) w = &vel->z->data[0][0][0];
    else {
      zero.assign(f_->size(), 0.0);
      w = zero.data();
    }
    check(scihpc_ls_set_velocity(h_, &vel->x->data[0][0][0], &vel->y->data[0][0][0], w));
  }
};

// weno_solver.h:17-48: f->fx, fy, fz from f->flux, upwinded by vel
class weno_solver : public level_set_solver {
 public:
  using level_set_solver::level_set_solver;
  void weno5_find_derivatives(scalar_data *f, vector_data *vel) {
    set_velocity(vel);
    check(scihpc_ls_weno5_derivatives(h_, &f->flux[0][0][0], &f->fx[0][0][0], &f->fy[0][0][0], &f->fz[0][0][0]));
  }
};

// runge_kutta.h:12-23: TVD-RK3 of d phi/dt = -u . grad(phi) with WENO5
// (default) or UCCD (set_scheme(SCIHPC_LS_UCCD): the reference's
// `convection`) derivatives of the identity flux and the solver's BCs after
// every stage; plus the reference's other right-hand sides run through
// runge_kutta: euler with mpls (mass correction) and tvd_rk3 with
// lsf_redistance_lambda (redistancing).  Each call copies phi in and out.
class runge_kutta : public level_set_solver {
 public:
  using level_set_solver::level_set_solver;
  void set_scheme(int scheme) { check(scihpc_ls_set_scheme(h_, scheme)); }
  void tvd_rk3(vector_data *vel, DataType dt, int steps = 1) {
    set_velocity(vel);
    check(scihpc_ls_set_phi(h_, &f_->data[0][0][0]));
    check(scihpc_ls_tvd_rk3(h_, dt, steps));
    check(scihpc_ls_get_phi(h_, &f_->data[0][0][0]));
  }
  // lsf_mass (wrapper_func.cpp:450-511)
  DataType lsf_mass(DataType ls_width, DataType density_ratio, bool positive_ref = true) {
    double m = 0.0;
    check(scihpc_ls_set_phi(h_, &f_->data[0][0][0]));
    check(scihpc_ls_mass(h_, ls_width, density_ratio, positive_ref ? 1 : 0, &m));
    return m;
  }
  // `count` x euler(phi, nvel, identity_flux, mpls) (source.cpp:109-146)
  void euler_mpls(DataType dt, DataType ls_width, DataType density_ratio, DataType mass0, int count = 1,
                  bool positive_ref = true) {
    check(scihpc_ls_set_phi(h_, &f_->data[0][0][0]));
    check(scihpc_ls_mpls(h_, dt, ls_width, density_ratio, positive_ref ? 1 : 0, mass0, count));
    check(scihpc_ls_get_phi(h_, &f_->data[0][0][0]));
  }
  // find_sign / find_delta (wrapper_func.cpp:20-40) of the current phi
  void find_sign(DataType ls_width) {
    check(scihpc_ls_set_phi(h_, &f_->data[0][0][0]));
    check(scihpc_ls_find_sign(h_, ls_width));
  }
  void find_delta(DataType ls_width) {
    check(scihpc_ls_set_phi(h_, &f_->data[0][0][0]));
    check(scihpc_ls_find_delta(h_, ls_width));
  }
  // `steps` x tvd_rk3(phi, nvel, identity_flux, lsf_redistance_lambda)
  // with the stored sign and delta (source.cpp:148-240)
  void tvd_rk3_redistance(DataType dt, DataType rdt, int steps = 1) {
    check(scihpc_ls_set_phi(h_, &f_->data[0][0][0]));
    check(scihpc_ls_redistance(h_, dt, rdt, steps));
    check(scihpc_ls_get_phi(h_, &f_->data[0][0][0]));
  }
};

// vtkWriter.h:15-35: legacy VTK snapshots, byte-identical to the
// reference's files; create(id) opens <filename><id>.vtk.  The device does
// the reorder and byte swap, a background thread writes the file, close()
// waits for it.  A 2-D vector_data of this facade has no z component: its
// VECTORS get zeros there, as the reference's zero-filled z field.
class vtkWriter {
 public:
  vtkWriter(structured_grid *geo, std::string filename) : grid(geo), filename(std::move(filename)) {
    const int n[3] = {geo->x_axis.n, geo->ndim > 1 ? geo->y_axis.n : 1, geo->ndim > 2 ? geo->z_axis.n : 1};
    const double org[3] = {geo->x_axis.start, geo->y_axis.start, geo->z_axis.start};
    const double sp[3] = {geo->dx, geo->dy, geo->dz};
    check(scihpc_vtk_create(geo->ndim, n[0], n[1], n[2], org, sp, this->filename.c_str(), &h_));
  }
  ~vtkWriter() { scihpc_vtk_destroy(h_); }
  vtkWriter(const vtkWriter &) = delete;
  vtkWriter &operator=(const vtkWriter &) = delete;

  structured_grid *grid;
  std::string filename;

  void create(unsigned int id) { check(scihpc_vtk_begin(h_, id)); }
  void add_scalar(scalar_data *data, const std::string &name) { add_scalar(data->data, data, name); }
  void add_scalar(DataType ***f, scalar_data *, const std::string &name) {
    check(scihpc_vtk_add_scalar(h_, name.c_str(), &f[0][0][0], 0));
  }
  void add_vector(vector_data *data, const std::string &name) {
    add_vector(data->x->data, data->y ? data->y->data : nullptr, data->z ? data->z->data : nullptr, data, name);
  }
  void add_vector(DataType ***u, DataType ***v, DataType ***w, vector_data *data, const std::string &name) {
    std::vector<DataType> zero;
    if (!u || !v || !w) zero.assign(data->x->size(), 0.0);
    auto ptr = [&](DataType ***f) -> const DataType * { return f ? &f[0][0][0] : zero.data(); };
    const DataType *pu = ptr(u), *pv = ptr(v), *pw = ptr(w);
    check(scihpc_vtk_add_vector(h_, name.c_str(), pu, pv, pw, 0));
  }
  void close() { check(scihpc_vtk_close(h_)); }

 private:
  scihpc_vtk *h_ = nullptr;
};

}  // namespace scihpc_gpu
