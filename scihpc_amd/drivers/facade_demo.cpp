// facade_demo.cpp -- SciHPC-style drivers written only against SciHPC's own
// driver API (include/scihpc/*.h, served by the MI355X library), plus the
// two additions of this build (multigrid::solve, multigrid::set_density):
// (1) the 2-D Dirichlet Poisson problem solved with the reference-faithful
//     V-cycle loop of benchmarks_2d_multigrid_Poisson_Dirichlet (residual()
//     loop around v_cycle(), operator chosen by init_DirichletBC());
// (2) the same system with the red-black mixed-precision solve();
// (3) a variable-density bubble solve through set_density();
// (4) a circle's level set moved by runge_kutta::tvd_rk3 with the
//     convection right-hand side, and weno_solver's derivatives of a linear
//     field;
// (5) the projection step (projection() on A(rho), then
//     fast_pressure_correction) through projection_method and wrappers.
// Prints "key value" lines for the tests.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "scihpc/multigrid.h"
#include "scihpc/projection_method.h"
#include "scihpc/runge_kutta.h"
#include "scihpc/structured_grid.h"
#include "scihpc/vtkWriter.h"
#include "scihpc/weno_solver.h"
#include "scihpc/wrapper.h"
#include "scihpc/wrapper_func.h"

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 64;
  // (1) + (2): -lap p = 2 pi^2 sin(pi x) sin(pi y), p = 0 on the walls
  auto geo = structured_grid(axis{0.0, 1.0, n}, axis{0.0, 1.0, n});
  auto phi = wrapper(true, &geo, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});
  auto mg = multigrid(phi.scalar, &geo);
  for (unsigned l = 0; l < mg.level_num; ++l) {
    mg.at[l]->no_compatibility = true;
    mg.at[l]->init_DirichletBC();
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      mg.at[0]->rhs[mg.at[0]->of(i, j)] = -2.0 * pi * pi * std::sin(pi * geo.xc[i]) * std::sin(pi * geo.yc[j]);
  int cycles = 0;
  double r = mg.at[0]->residual();
  const double r0 = r;
  while (r > 1e-8 * r0 && cycles < 200) {
    mg.v_cycle();
    r = mg.at[0]->residual();
    ++cycles;
  }
  auto error = [&](const VectorX &s) {
    double e = 0.0, m = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double ex = std::sin(pi * geo.xc[i]) * std::sin(pi * geo.yc[j]);
        e += (s[mg.at[0]->of(i, j)] - ex) * (s[mg.at[0]->of(i, j)] - ex);
        m += ex * ex;
      }
    return std::sqrt(e / m);
  };
  std::printf("levels %u\n", mg.level_num);
  std::printf("vcycle_cycles %d\n", cycles);
  std::printf("vcycle_rel_residual %.6e\n", r / r0);
  std::printf("vcycle_error %.6e\n", error(mg.at[0]->sol));
  VectorX ref_sol = mg.at[0]->sol;
  mg.at[0]->sol.setZero();
  scihpc_solve_stats st = mg.solve(1e-10, SCIHPC_PREC_MIXED, 1);
  double dmax = 0.0;
  for (long q = 0; q < ref_sol.size(); ++q) dmax = std::fmax(dmax, std::fabs(ref_sol[q] - mg.at[0]->sol[q]));
  std::printf("solve_cycles %d\n", st.cycles);
  std::printf("solve_rel_residual %.6e\n", st.rel_residual);
  std::printf("solve_error %.6e\n", error(mg.at[0]->sol));
  std::printf("solve_vs_vcycle_maxdiff %.6e\n", dmax);

  // (3) bubble: rho 1000 outside / 1 inside r = 0.25 of a 1 x 2 box
  auto geo2 = structured_grid(axis{0.0, 1.0, n}, axis{0.0, 2.0, 2 * n});
  auto lsf = wrapper(true, &geo2, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});
  auto param = new problem_parameters();
  param->ls_width = 1.5 * geo2.dx;
  param->density_ratio = 1.0 / 1000.0;
  param->positive_ref = true;
  auto solvers = SharedSolvers_alloc(lsf.scalar, &geo2);
  shared_solvers_mg_init_Neumann(solvers);
  auto dummy = dummy_data_alloc(lsf.scalar);
  lsf.link_params(param);
  lsf.link_solvers(solvers);
  lsf.link_dummy(dummy);
  for (int I = 0; I < lsf.scalar->Nx; ++I)
    for (int J = 0; J < lsf.scalar->Ny; ++J)
      lsf.scalar->data[I][J][0] =
          0.25 - std::hypot(geo2.dx * (I - 3 + 0.5) - 0.5, geo2.dy * (J - 3 + 0.5) - 0.5);
  find_density(&lsf);  // 1 inside (phi > 0), 1/1000 outside: the reference's nondimensional density
  auto mg2 = multigrid(lsf.scalar, &geo2);
  scalar_data rho(n, 2 * n);
  for (int I = 0; I < rho.Nx; ++I)
    for (int J = 0; J < rho.Ny; ++J) rho.data[I][J][0] = 1000.0 * dummy->density[I][J][0];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j)
      mg2.at[0]->rhs[mg2.at[0]->of(i, j)] = std::cos(pi * geo2.xc[i]) * std::cos(pi * geo2.yc[j]);
  mg2.set_density(&rho);
  scihpc_solve_stats s2 = mg2.solve(1e-8, SCIHPC_PREC_MIXED, 1);
  std::printf("bubble_cycles %d\n", s2.cycles);
  std::printf("bubble_rel_residual %.6e\n", s2.rel_residual);
  std::printf("bubble_mass %.6e\n", lsf_mass(&lsf));

  // (4) level set of a circle moved by u = (1, 0.5) for 10 steps with the
  // convection right-hand side (UCCD, source.cpp:8-40)
  auto ls = wrapper(true, &geo, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});
  auto nvel = wrapper(false, &geo, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});
  auto p4 = new problem_parameters();
  p4->dt = 0.25 * geo.dx;
  ls.link_params(p4);
  nvel.link_params(p4);
  for (int I = 0; I < ls.scalar->Nx; ++I)
    for (int J = 0; J < ls.scalar->Ny; ++J) {
      const double xc = geo.dx * (I - ls.scalar->ghc + 0.5), yc = geo.dy * (J - ls.scalar->ghc + 0.5);
      ls.scalar->data[I][J][0] = 0.2 - std::hypot(xc - 0.4, yc - 0.4);
      nvel.vector->x.data[I][J][0] = 1.0;
      nvel.vector->y.data[I][J][0] = 0.5;
    }
  auto rk = runge_kutta(ls.scalar->Nx, ls.scalar->Ny, ls.scalar->Nz);
  for (int s = 0; s < 10; ++s) rk.tvd_rk3(&ls, &nvel, &identity_flux, &convection);
  const double dt = p4->dt;
  double e = 0.0;
  int cnt = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const indices q = ls.scalar->index_mapping(i + 1, j + 1, 1);
      const double ex = 0.2 - std::hypot(geo.xc[i] - 0.4 - 10 * dt, geo.yc[j] - 0.4 - 5 * dt);
      if (std::fabs(ex) < 0.1) {
        e = std::fmax(e, std::fabs(ls.scalar->data[q.i][q.j][q.k] - ex));
        ++cnt;
      }
    }
  std::printf("levelset_max_error_near_interface %.6e\n", e);
  std::printf("levelset_cells %d\n", cnt);
  // weno_solver on a linear flux 3x - 2y: the derivatives are exact
  auto weno = weno_solver(ls.scalar, &geo);
  for (int I = 0; I < ls.scalar->Nx; ++I)
    for (int J = 0; J < ls.scalar->Ny; ++J)
      ls.scalar->flux[I][J][0] = 3.0 * geo.dx * (I - 3 + 0.5) - 2.0 * geo.dy * (J - 3 + 0.5);
  weno.weno5_find_derivatives(ls.scalar, nvel.vector);
  double we = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const indices q = ls.scalar->index_mapping(i + 1, j + 1, 1);
      we = std::fmax(we, std::fabs(ls.scalar->fx[q.i][q.j][0] - 3.0) + std::fabs(ls.scalar->fy[q.i][q.j][0] + 2.0));
    }
  std::printf("weno_linear_error %.6e\n", we);
  if (argc > 2) {
    // snapshot of the transported level set and its velocity, as the
    // reference's drivers write them (benchmarks_2d_dambreak.cpp:94-98)
    auto vtk = vtkWriter(&geo, argv[2]);
    vtk.create(10);
    vtk.add_scalar(ls.scalar, "phi");
    vtk.add_vector(nvel.vector, "nvel");
    vtk.close();
    std::printf("vtk_file %s10.vtk\n", argv[2]);
  }

  // (5) projection steps on the bubble grid: the A(rho) projection and the
  // FPC step; both remove the divergence of a wall-consistent field
  auto pres = wrapper(true, &geo2, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});
  auto vel2 = wrapper(false, &geo2, bc_info{NO_SLIP}, bc_info{NO_SLIP}, bc_info{NO_SLIP}, bc_info{NO_SLIP});
  for (wrapper *w : {&pres, &vel2}) {
    w->link_params(param);
    w->link_solvers(solvers);
    w->link_dummy(dummy);
  }
  param->dt = param->dt_old = param->dt_old2 = 0.01;
  param->ppe_tol = 1e-9;
  const int g = pres.scalar->ghc;
  for (int I = g; I < g + n; ++I)
    for (int J = g; J < g + 2 * n; ++J) {
      if (I < g + n - 1) vel2.vector->x.data[I][J][0] = 0.5 * std::sin(7.0 * I + J);
      if (J < g + 2 * n - 1) vel2.vector->y.data[I][J][0] = 0.5 * std::cos(3.0 * I - 2.0 * J);
    }
  vel2.apply_vel_bc();
  auto pm = projection_method(lsf.scalar);
  const double d_before = divergence(&vel2);
  pm.projection(&pres, &lsf, &vel2);
  std::printf("projection_cycles %d\n", pm.last_cycles);
  std::printf("projection_div_ratio %.6e\n", divergence(&vel2) / d_before);
  for (int I = g; I < g + n - 1; ++I)
    for (int J = g; J < g + 2 * n; ++J) vel2.vector->x.data[I][J][0] += 0.1 * std::sin(5.0 * I - 3.0 * J);
  pres.scalar->store();
  pres.scalar->store();
  const double d_fpc = divergence(&vel2);
  pm.fast_pressure_correction(&pres, &lsf, &vel2);
  std::printf("fpc_cycles %d\n", pm.last_cycles);
  std::printf("fpc_div_ratio %.6e\n", divergence(&vel2) / d_fpc);
  return 0;
}
