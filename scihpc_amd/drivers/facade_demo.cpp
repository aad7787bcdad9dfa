// facade_demo.cpp -- SciHPC-style drivers written against scihpc_facade.hpp:
// (1) the 2-D Dirichlet Poisson problem solved with the reference-faithful
//     V-cycle loop of benchmarks_2d_multigrid_Poisson_Dirichlet (residual()
//     loop around v_cycle(), operator chosen by init_DirichletBC()),
// (2) the same system solved with the red-black mixed-precision solve(),
// (3) a variable-density bubble solve through set_density(),
// (4) WENO5 / TVD-RK3 transport of a circle's level set.
// Prints "key value" lines for the tests.
#include <cmath>
#include <cstdio>

#include "scihpc_facade.hpp"

using namespace scihpc_gpu;

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 64;
  const double pi = 3.14159265358979323846;
  // (1) + (2): -lap p = 2 pi^2 sin(pi x) sin(pi y), p = 0 on the walls
  structured_grid geo(axis{0.0, 1.0, n}, axis{0.0, 1.0, n});
  scalar_data phi(n, n);
  multigrid mg(&phi, &geo);
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

  // (3) bubble: rho 1000 outside / 1 inside r = 0.25, b = div of a smooth field
  structured_grid geo2(axis{0.0, 1.0, n}, axis{0.0, 2.0, 2 * n});
  scalar_data rho(n, 2 * n);
  multigrid mg2(&rho, &geo2);
  const double w = 1.5 * geo2.dx;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) {
      const double d = 0.25 - std::hypot(geo2.xc[i] - 0.5, geo2.yc[j] - 0.5);
      double hv = d < -w ? 0.0 : d > w ? 1.0 : 0.5 * (1.0 + d / w + std::sin(pi * d / w) / pi);
      const indices q = rho.index_mapping(i + 1, j + 1, 1);
      rho.data[q.i][q.j][q.k] = hv + (1.0 - hv) * 1000.0;
      mg2.at[0]->rhs[mg2.at[0]->of(i, j)] = std::cos(pi * geo2.xc[i]) * std::cos(pi * geo2.yc[j]);
    }
  // ghost cells of rho are never read by the operator
  mg2.set_density(&rho);
  scihpc_solve_stats s2 = mg2.solve(1e-8, SCIHPC_PREC_MIXED, 1);
  std::printf("bubble_cycles %d\n", s2.cycles);
  std::printf("bubble_rel_residual %.6e\n", s2.rel_residual);

  // (4) level set of a circle moved by u = (1, 0.5) for 10 steps
  scalar_data ls(n, n);
  vector_data vel(n, n);
  for (int I = 0; I < ls.Nx; ++I)
    for (int J = 0; J < ls.Ny; ++J) {
      const double xc = geo.dx * (I - ls.ghc + 0.5), yc = geo.dy * (J - ls.ghc + 0.5);
      ls.data[I][J][0] = 0.2 - std::hypot(xc - 0.4, yc - 0.4);
      vel.x->data[I][J][0] = 1.0;
      vel.y->data[I][J][0] = 0.5;
    }
  bc_info faces[6] = {{NEUMANN}, {NEUMANN}, {NEUMANN}, {NEUMANN}, {NEUMANN}, {NEUMANN}};
  runge_kutta rk(&ls, &geo, faces);
  const double dt = 0.25 * geo.dx;
  rk.tvd_rk3(&vel, dt, 10);
  // centre of the zero contour moves by 10 dt (1, 0.5): compare phi with the
  // shifted exact signed distance on the interior
  double e = 0.0;
  int cnt = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const indices q = ls.index_mapping(i + 1, j + 1, 1);
      const double ex = 0.2 - std::hypot(geo.xc[i] - 0.4 - 10 * dt, geo.yc[j] - 0.4 - 5 * dt);
      if (std::fabs(ex) < 0.1) {
        e = std::fmax(e, std::fabs(ls.data[q.i][q.j][q.k] - ex));
        ++cnt;
      }
    }
  std::printf("levelset_max_error_near_interface %.6e\n", e);
  std::printf("levelset_cells %d\n", cnt);
  if (argc > 2) {
    // snapshot of the transported level set and its velocity, as the
    // reference's drivers write them (benchmarks_2d_dambreak.cpp:94-98)
    vtkWriter vtk(&geo, argv[2]);
    vtk.create(10);
    vtk.add_scalar(&ls, "phi");
    vtk.add_vector(&vel, "nvel");
    vtk.close();
    std::printf("vtk_file %s10.vtk\n", argv[2]);
  }

  // (5) projection steps on the bubble grid: FPC (fast_pressure_correction
  // with the reference-faithful full cycles) and the A(rho) projection; both
  // remove most of the divergence of a wall-consistent random field
  scalar_data pres(n, 2 * n);
  vector_data vel2(n, 2 * n);
  for (int I = 0; I < pres.Nx; ++I)
    for (int J = 0; J < pres.Ny; ++J) {
      vel2.x->data[I][J][0] = std::sin(3.0 * I) * std::cos(2.0 * J);
      vel2.y->data[I][J][0] = std::cos(5.0 * I) * std::sin(1.0 * J);
    }
  bc_info walls[6] = {{NO_SLIP}, {NO_SLIP}, {NO_SLIP}, {NO_SLIP}, {NO_SLIP}, {NO_SLIP}};
  auto divergence = [&](const vector_data &vv) {
    double s = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < 2 * n; ++j) {
        const indices q = pres.index_mapping(i + 1, j + 1, 1);
        const double d = (vv.x->data[q.i][q.j][0] - vv.x->data[q.i - 1][q.j][0]) / geo2.dx +
                         (vv.y->data[q.i][q.j][0] - vv.y->data[q.i][q.j - 1][0]) / geo2.dy;
        s += d * d;
      }
    return std::sqrt(s);
  };
  projection_method pm(&pres, &geo2, walls);
  // walls as the velocity BCs leave them: the step's own BCs on a zero update
  scihpc_solve_stats s0 = pm.projection(mg2, &pres, &rho, &vel2, 1e-3, 1e-3);
  (void)s0;
  // perturb interior faces only: wall faces stay zero, so div u sums to zero
  const int g = pres.ghc;
  for (int I = g; I < g + n; ++I)
    for (int J = g; J < g + 2 * n; ++J) {
      if (I < g + n - 1) vel2.x->data[I][J][0] += 0.5 * std::sin(7.0 * I + J);
      if (J < g + 2 * n - 1) vel2.y->data[I][J][0] += 0.5 * std::cos(3.0 * I - 2.0 * J);
    }
  const double d_before = divergence(vel2);
  scihpc_solve_stats s3 = pm.projection(mg2, &pres, &rho, &vel2, 0.01, 1e-10);
  std::printf("projection_rel_residual %.6e\n", s3.rel_residual);
  std::printf("projection_div_ratio %.6e\n", divergence(vel2) / d_before);
  multigrid mg3(&pres, &geo2);  // constant-coefficient Neumann hierarchy for FPC
  problem_parameters prm;
  prm.dt = 0.01;
  prm.dt_old = 0.01;
  prm.dt_old2 = 0.01;
  prm.density_ratio = 1.0 / 1000.0;
  prm.ppe_tol = 1e-9;
  pres.store();
  pres.store();
  const int fc = pm.fast_pressure_correction(mg3, &pres, &rho, &vel2, prm);
  std::printf("fpc_cycles %d\n", fc);
  return 0;
}
