// dambreak_demo.cpp -- the 2-D dam break of SciHPC's
// benchmarks_2d_dambreak.cpp written against include/scihpc_facade.hpp, so
// every step runs on the GPU: level-set transport (TVD-RK3), the momentum
// predictor, the pressure step (projection() for the first 5 steps, then the
// FPC step) and node velocities, with legacy-VTK snapshots.
//
// As the reference driver: the level set is advected with the UCCD
// convection term, corrected by 5 mpls Euler steps per time step, and
// redistanced every 10 steps (lsf_redistance_lambda with the sign found at
// the start, as the reference's loop leaves it).  One difference: the level
// set starts as the exact signed distance to the water column (the
// reference redistances a +-1 indicator first).
//
//   dambreak_demo <n> <steps> [vtk prefix] [plain]   (n x n cells on [0,5]^2)
// `plain` (any 4th argument) turns the interface maintenance off: WENO5
// convection only.
// prints key/value lines: the time, the water volume, the largest
// velocity, the divergence after the pressure step and the front position.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "scihpc_facade.hpp"

using namespace scihpc_gpu;

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 64;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 100;
  const char *prefix = argc > 3 && argv[3][0] ? argv[3] : nullptr;
  const bool maintain = argc <= 4;

  structured_grid geo(axis{0.0, 5.0, n}, axis{0.0, 5.0, n});
  scalar_data phi(n, n), pressure(n, n), density(n, n);
  vector_data vel(n, n), nvel(n, n);
  const bc_info walls[6] = {{NO_SLIP}, {NO_SLIP}, {NO_SLIP}, {NO_SLIP}, {NO_SLIP}, {NO_SLIP}};
  const bc_info neumann[6] = {{NEUMANN}, {NEUMANN}, {NEUMANN}, {NEUMANN}, {NEUMANN}, {NEUMANN}};

  // water column [0,1] x [0,1]: signed distance, positive inside
  for (int I = 0; I < phi.Nx; ++I)
    for (int J = 0; J < phi.Ny; ++J) {
      const double x = geo.dx * (I - phi.ghc + 0.5), y = geo.dy * (J - phi.ghc + 0.5);
      const double dx = x - 1.0, dy = y - 1.0;
      double d;
      if (dx <= 0 && dy <= 0) d = -std::fmax(dx, dy);
      else d = -std::hypot(std::fmax(dx, 0.0), std::fmax(dy, 0.0));
      phi.data[I][J][0] = d;
    }

  // set_air_water(0.0571) (global.cpp:30-49, :73-76) and the driver's settings
  const double rho1 = 1000.0, rho2 = 1.226, mu1 = 1.137e-3, mu2 = 1.78e-5, g = 9.81, L = 0.0571;
  const double Uc = std::sqrt(g * L);
  problem_parameters prm;
  prm.density_ratio = rho2 / rho1;
  prm.viscosity_ratio = mu2 / mu1;
  prm.Reynolds_number = rho1 * Uc * L / mu1;
  prm.Froude_number = Uc / std::sqrt(g * L);
  prm.Weber_number = -1.0;
  prm.ls_width = 1.5 * geo.h;
  prm.ppe_tol = 1e-8;
  prm.ppe_initer = 1;
  const double dt = 0.01 * geo.h;  // find_dt with min_CFL = max_CFL = 0.01
  prm.dt = prm.dt_old = prm.dt_old2 = dt;

  runge_kutta rk(&phi, &geo, neumann);
  if (maintain) rk.set_scheme(SCIHPC_LS_UCCD);
  const double rdt = 0.25 * geo.h;  // param->rdt of the driver
  projection_method flow(&pressure, &geo, walls);
  multigrid mg_const(&pressure, &geo);  // FPC: constant-coefficient Neumann hierarchy
  multigrid mg_rho(&pressure, &geo);    // projection(): A(rho)
  vtkWriter vtk(&geo, prefix ? prefix : "dambreak");

  auto volume = [&]() {
    double s = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const indices q = phi.index_mapping(i + 1, j + 1, 1);
        s += phi.data[q.i][q.j][0] > 0.0 ? geo.dv : 0.0;
      }
    return s;
  };
  auto max_div = [&]() {
    double m = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const indices q = phi.index_mapping(i + 1, j + 1, 1);
        const double d = (vel.x->data[q.i][q.j][0] - vel.x->data[q.i - 1][q.j][0]) / geo.dx +
                         (vel.y->data[q.i][q.j][0] - vel.y->data[q.i][q.j - 1][0]) / geo.dy;
        m = std::fmax(m, std::fabs(d));
      }
    return m;
  };
  auto front = [&]() {  // right-most wet cell centre on the bottom row
    const int j = 0;
    double f = 0.0;
    for (int i = 0; i < n; ++i) {
      const indices q = phi.index_mapping(i + 1, j + 1, 1);
      if (phi.data[q.i][q.j][0] > 0.0) f = geo.xc[i];
    }
    return f;
  };
  auto max_speed = [&]() {
    double m = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const indices q = phi.index_mapping(i + 1, j + 1, 1);
        m = std::fmax(m, std::hypot(nvel.x->data[q.i][q.j][0], nvel.y->data[q.i][q.j][0]));
      }
    return m;
  };

  const double v0 = volume(), f0 = front();
  const double mass0 = maintain ? rk.lsf_mass(prm.ls_width, prm.density_ratio) : 0.0;
  if (maintain) rk.find_sign(prm.ls_width);
  std::printf("volume0 %.6e\n", v0);
  std::printf("front0 %.6e\n", f0);
  if (prefix) {
    vtk.create(0);
    vtk.add_scalar(&phi, "phi");
    vtk.add_scalar(&pressure, "pressure");
    vtk.add_vector(&nvel, "nvel");
    vtk.close();
  }
  double t = 0.0;
  int cycles = 0;
  for (int s = 1; s <= steps; ++s) {
    t += dt;
    prm.iter = s;
    rk.tvd_rk3(&nvel, dt, 1);
    if (maintain) rk.euler_mpls(dt, prm.ls_width, prm.density_ratio, mass0, 5);
    cycles = flow.ab_solve_sec(mg_const, mg_rho, &vel, &nvel, &pressure, &phi, &density, prm);
    if (maintain && s % 10 == 0) {
      // the reference's redistancing loop (benchmarks_2d_dambreak.cpp:125-131):
      // delta as find_source_sec left it, the initial sign
      rk.find_delta(prm.ls_width);
      int instep = 0;
      int n_re = 0;
      while (++instep * rdt < 5.0 * prm.ls_width) ++n_re;
      rk.tvd_rk3_redistance(dt, rdt, n_re);
    }
  }
  if (maintain) std::printf("mass_ratio %.6e\n", rk.lsf_mass(prm.ls_width, prm.density_ratio) / mass0);
  std::printf("time %.6e\n", t);
  std::printf("steps %d\n", steps);
  std::printf("last_pressure_cycles %d\n", cycles);
  std::printf("volume_ratio %.6e\n", volume() / v0);
  std::printf("max_divergence %.6e\n", max_div());
  std::printf("max_speed %.6e\n", max_speed());
  std::printf("front %.6e\n", front());
  if (prefix) {
    vtk.create(1);
    vtk.add_scalar(&phi, "phi");
    vtk.add_scalar(&pressure, "pressure");
    vtk.add_vector(&nvel, "nvel");
    vtk.close();
    std::printf("vtk_file %s1.vtk\n", prefix);
  }
  return 0;
}
