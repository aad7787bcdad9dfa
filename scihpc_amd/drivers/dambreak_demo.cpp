// dambreak_demo.cpp -- SciHPC's 2-D dam break (benchmarks_2d_dambreak.cpp)
// written only against SciHPC's own driver API (include/scihpc/*.h, served
// by the MI355X library): the same objects, the same calls in the same
// order -- wrapper, runge_kutta::tvd_rk3 with convection, euler with mpls,
// projection_method::ab_solve_sec, the redistancing every 10 steps,
// vtkWriter -- with the grid, the box, the water column, the time step and
// the snapshots taken from the command line:
//
//   dambreak_demo <nx> <ny> <lx> <ly> <col_w> <col_h> <steps> <cfl> <init>
//                 [vtk prefix] [snapshot every k steps]
//
// init "indicator": phi = +-1 in/out of the column, then find_sign,
// stabilized_upon_gradient and redistancing until l2norm < 1e-6 (the
// reference's initialisation); init "distance": the exact signed distance
// to the column's free faces.  dt = min(stable_CFL, cfl) * h: find_dt's own
// stability bound (wrapper_func.cpp:335-383) caps the step, min_CFL =
// min(cfl, 0.01) (the reference's driver pins both at 0.01).  A step whose
// divergence or level-set mass is not finite stops the run with status 3
// (after printing stopped_at and the snapshot count).  Every k steps (k > 0) a snapshot <prefix><id>.vtk
// holds phi, pressure, density, the pressure-step right-hand side
// rhs = div(u*) / dt of the intermediate velocity, and nvel -- the (rhs,
// rho, p) triples the learned preconditioner trains on.  Prints key/value
// lines on stdout and a progress line every 100 steps on stderr.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "scihpc/bc_factory.h"
#include "scihpc/flux.h"
#include "scihpc/godunov_gradient.h"
#include "scihpc/projection_method.h"
#include "scihpc/runge_kutta.h"
#include "scihpc/structured_grid.h"
#include "scihpc/vtkWriter.h"
#include "scihpc/wrapper.h"
#include "scihpc/wrapper_func.h"

int main(int argc, char **argv) {
  if (argc < 10) {
    std::fprintf(stderr, "usage: %s nx ny lx ly col_w col_h steps cfl indicator|distance [prefix] [every]\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]);
  const double lx = std::atof(argv[3]), ly = std::atof(argv[4]);
  const double cw = std::atof(argv[5]), ch = std::atof(argv[6]);
  const int steps = std::atoi(argv[7]);
  const double cfl = std::atof(argv[8]);
  const bool indicator = std::strcmp(argv[9], "indicator") == 0;
  const char *prefix = argc > 10 && argv[10][0] ? argv[10] : nullptr;
  const int every = argc > 11 ? std::atoi(argv[11]) : 0;

  auto geo = structured_grid(axis{0.0, lx, nx}, axis{0.0, ly, ny});

  auto phi = wrapper(true, &geo, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});
  auto vel = wrapper(false, &geo, bc_info{NO_SLIP}, bc_info{NO_SLIP}, bc_info{NO_SLIP}, bc_info{NO_SLIP});
  auto nvel = wrapper(false, &geo, bc_info{NO_SLIP}, bc_info{NO_SLIP}, bc_info{NO_SLIP}, bc_info{NO_SLIP});
  auto pressure = wrapper(true, &geo, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN}, bc_info{NEUMANN});

  auto solver = runge_kutta(phi.scalar->Nx, phi.scalar->Ny, phi.scalar->Nz);
  auto flow_solver = projection_method(phi.scalar);

  auto param = set_air_water(0.0571);
  auto deri_solvers = SharedSolvers_alloc(phi.scalar, &geo);
  shared_solvers_mg_init_Neumann(deri_solvers);
  auto dummy = dummy_data_alloc(phi.scalar);

  for (int i = 0; i < phi.scalar->nx; ++i) {
    for (int j = 0; j < phi.scalar->ny; ++j) {
      auto index = phi.scalar->index_mapping(i + 1, j + 1, 1);
      const double x = geo.xc[i], y = geo.yc[j];
      double d;
      if (indicator) {
        d = (x < cw && y < ch) ? 1.0 : -1.0;
      } else {
        // signed distance to the free faces x = cw (and y = ch below the lid)
        const double ex = x - cw, ey = ch < ly ? y - ch : -1e30;
        if (ex <= 0 && ey <= 0) d = -std::fmax(ex, ey);
        else d = -std::hypot(std::fmax(ex, 0.0), std::fmax(ey, 0.0));
      }
      phi.scalar->data[index.i][index.j][index.k] = d;
    }
  }
  phi.apply_scalar_bc();

  for (wrapper *w : {&phi, &pressure, &vel, &nvel}) {
    w->link_params(param);
    w->link_solvers(deri_solvers);
    w->link_dummy(dummy);
  }

  param->ls_width = 1.5 * geo.h;
  param->rdt = 0.25 * geo.h;
  param->max_CFL = cfl;
  param->min_CFL = std::fmin(cfl, 0.01);
  param->Weber_number = -1.0;
  param->ppe_tol = 1e-8;
  param->ppe_initer = 1;

  std::printf("reynolds %.6e\nfroude %.6e\n", param->Reynolds_number, param->Froude_number);

  // the reference's initialisation of phi (benchmarks_2d_dambreak.cpp:77-83)
  int step = 0;
  find_sign(&phi);
  if (indicator) {
    stabilized_upon_gradient(&phi);
    do {
      store_tmp(&phi);
      solver.tvd_rk3(&phi, &vel, identity_flux, lsf_redistance_lambda);
    } while (++step * param->rdt < 5.0 and l2norm(&phi) > 1e-6);
  }
  std::printf("init_redistance_steps %d\n", step);

  flow_solver.find_source(&vel, &nvel, &phi);
  param->lsf_mass0 = lsf_mass(&phi);

  // snapshot buffers: density and the pressure-step right-hand side
  scalar_data rhs(nx, ny);
  auto snapshot = [&](vtkWriter &vtk, unsigned id) {
    vtk.create(id);
    vtk.add_scalar(phi.scalar, "phi");
    vtk.add_scalar(pressure.scalar, "pressure");
    vtk.add_scalar(dummy->density, phi.scalar, "density");
    vtk.add_scalar(&rhs, "rhs");
    vtk.add_vector(nvel.vector, "nvel");
    vtk.close();
  };
  auto front = [&]() {  // right-most wet cell centre on the bottom row
    double f = 0.0;
    for (int i = 0; i < nx; ++i) {
      const indices q = phi.scalar->index_mapping(i + 1, 1, 1);
      if (phi.scalar->data[q.i][q.j][0] > 0.0) f = geo.xc[i];
    }
    return f;
  };
  auto max_speed = [&]() {
    double m = 0.0;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const indices q = phi.scalar->index_mapping(i + 1, j + 1, 1);
        m = std::fmax(m, std::hypot(nvel.vector->x.data[q.i][q.j][0], nvel.vector->y.data[q.i][q.j][0]));
      }
    return m;
  };
  // rhs = div(u*) / dt on the interior (projection_method.cpp:546-568)
  auto record_rhs = [&]() {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const indices q = phi.scalar->index_mapping(i + 1, j + 1, 1);
        const auto &u = vel.vector->x.data, &v = vel.vector->y.data;
        rhs.data[q.i][q.j][0] = ((u[q.i][q.j][0] - u[q.i - 1][q.j][0]) / geo.dx +
                                 (v[q.i][q.j][0] - v[q.i][q.j - 1][0]) / geo.dy) / param->dt;
      }
  };

  std::unique_ptr<vtkWriter> vtk;
  if (prefix) {
    vtk.reset(new vtkWriter(&geo, prefix));
    snapshot(*vtk, 0);
  }
  std::printf("front0 %.6e\n", front());
  std::printf("mass0 %.6e\n", param->lsf_mass0);

  const auto t0 = std::chrono::steady_clock::now();
  unsigned snap_id = 1;
  int max_cycles = 0;
  double max_div = 0.0;
  for (int s = 1; s <= steps; ++s) {
    find_dt(&vel);
    param->t += param->dt;
    param->iter++;

    solver.tvd_rk3(&phi, &nvel, &identity_flux, &convection);
    for (int cnt = 0; cnt < 5; ++cnt) solver.euler(&phi, &nvel, &identity_flux, &mpls);

    const bool snap = prefix && every > 0 && param->iter % every == 0;
    if (snap) {
      // ab_solve_sec in its parts, to record the right-hand side the
      // pressure step sees (projection_method.cpp:779-800)
      pressure.scalar->store();
      flow_solver.find_source_sec(&vel, &nvel, &phi);
      flow_solver.find_intermediate_velocity(&vel);
      record_rhs();
      for (int cnt = 0; cnt < param->ppe_initer; ++cnt) {
        if (param->iter < 5) flow_solver.projection(&pressure, &phi, &vel);
        else flow_solver.fast_pressure_correction(&pressure, &phi, &vel);
      }
      node_from_face(&vel, &nvel);
    } else {
      flow_solver.ab_solve_sec(&vel, &nvel, &pressure, &phi);
    }
    max_cycles = std::max(max_cycles, flow_solver.last_cycles);
    const double div = divergence(&vel);
    const double mass = lsf_mass(&phi);
    if (!std::isfinite(div) || !std::isfinite(mass)) {
      std::fprintf(stderr, "step %d: divergence %g, level-set mass %g; stopping\n", s, div, mass);
      std::printf("stopped_at %d\nsnapshots %u\n", s, prefix ? snap_id : 0u);
      return 3;
    }
    max_div = std::fmax(max_div, div);

    if (param->iter % 10 == 0) {
      int instep = 0;
      while (++instep * param->rdt < 5.0 * param->ls_width) {
        if (instep == 0) find_sign(&phi);  // never reached, as in the reference
        solver.tvd_rk3(&phi, &nvel, identity_flux, lsf_redistance_lambda);
      }
    }
    if (snap) snapshot(*vtk, snap_id++);
    if (s % 100 == 0) {  // progress on stderr (stdout holds the key/value results)
      std::fprintf(stderr, "step %d t %.4f dt/h %.4f pressure_cycles %d front %.4f seconds %.1f\n", s, param->t,
                   param->dt / geo.h, flow_solver.last_cycles, front(),
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      std::fflush(stderr);
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  std::printf("time %.6e\n", param->t);
  std::printf("steps %d\n", steps);
  std::printf("dt %.6e\n", param->dt);
  std::printf("max_pressure_cycles %d\n", max_cycles);
  std::printf("last_pressure_cycles %d\n", flow_solver.last_cycles);
  std::printf("mass_ratio %.6e\n", lsf_mass(&phi) / param->lsf_mass0);
  std::printf("max_divergence %.6e\n", max_div);
  std::printf("final_divergence %.6e\n", divergence(&vel));
  std::printf("max_speed %.6e\n", max_speed());
  std::printf("front %.6e\n", front());
  std::printf("snapshots %u\n", prefix ? snap_id : 0u);
  std::printf("seconds %.3f\n", secs);
  return 0;
}
