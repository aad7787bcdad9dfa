// advect.hip -- level-set advection for gfx950 behind the C ABI of
// include/scihpc_hip.h: WENO5-Z derivatives (weno_solver.cpp:16-278),
// convection RHS s = -u.grad(phi) (source.cpp:8-40), TVD-RK3 in the
// reference's cumulative form (runge_kutta.cpp:13-49) and the bc_factory
// ghost fills (bc_factory.cpp:61-375 with neumann.cpp, dirichlet.cpp,
// periodic.cpp, no_slip.cpp, slip.cpp).
//
// Storage: the reference's padded scalar_data layout made contiguous
// (data[i][j][k], k fastest, 3 ghost cells on every axis < ndim).  One
// fused kernel per RK stage computes the WENO-Z face value on the upwind
// side only, the convection RHS and the stage update into a second
// buffer; one small kernel per axis then fills that buffer's ghosts in the
// reference's order (x, y, z; right side before left side).  Corner ghosts
// are never written, as in the reference.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace scihpc {

constexpr int kGhc = 3;
constexpr int kLsBlock = 256;

struct Grid {
  int ndim;
  int n[3], N[3], off[3];
  long s[3];
  long total;     // padded cells
  long interior;  // interior cells
  double h[3];
};

static Grid make_grid(int ndim, int nx, int ny, int nz, double dx, double dy, double dz) {
  Grid g{};
  g.ndim = ndim;
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    if (a < ndim) {
      g.n[a] = n[a];
      g.N[a] = n[a] + 2 * kGhc;
      g.off[a] = kGhc;
    } else {
      g.n[a] = 1;
      g.N[a] = 1;
      g.off[a] = 0;
    }
  }
  g.s[2] = 1;
  g.s[1] = g.N[2];
  g.s[0] = (long)g.N[1] * g.N[2];
  g.total = (long)g.N[0] * g.N[1] * g.N[2];
  g.interior = (long)g.n[0] * g.n[1] * g.n[2];
  g.h[0] = dx;
  g.h[1] = dy;
  g.h[2] = dz;
  return g;
}

// padded index of interior cell number t (last axis fastest)
__device__ __forceinline__ long interior_index(const Grid &g, long t) {
  const int k = (int)(t % g.n[2]);
  const long r = t / g.n[2];
  const int j = (int)(r % g.n[1]);
  const int i = (int)(r / g.n[1]);
  return (long)(i + g.off[0]) * g.s[0] + (long)(j + g.off[1]) * g.s[1] + (k + g.off[2]);
}

// ------------------------------------------------------------- WENO-Z
// Face value at I+1/2 from the five values a..e (left-biased, centred on c:
// interpolation_m/indicators_m) with the WENO-Z weights of
// weno_solver.cpp:50-66 (epsilon = 1e-8, global.h:14).
__device__ __forceinline__ double weno_face(double a, double b, double c, double d, double e,
                                            bool left_biased) {
  double q0, q1, q2, b0, b1, b2;
  if (left_biased) {
    q2 = (-e + 5.0 * d + 2.0 * c) / 6.0;
    q1 = (2.0 * d + 5.0 * c - b) / 6.0;
    q0 = (11.0 * c - 7.0 * b + 2.0 * a) / 6.0;
    b0 = 13.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) + 3.0 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
    b1 = 13.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 3.0 * (b - d) * (b - d);
    b2 = 13.0 * (c - 2.0 * d + e) * (c - 2.0 * d + e) + 3.0 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e);
  } else {
    // right-biased: a..e = f[I-1..I+3], centre c = f[I+1]
    q2 = (-a + 5.0 * b + 2.0 * c) / 6.0;
    q1 = (2.0 * b + 5.0 * c - d) / 6.0;
    q0 = (11.0 * c - 7.0 * d + 2.0 * e) / 6.0;
    b0 = 13.0 * (c - 2.0 * d + e) * (c - 2.0 * d + e) + 3.0 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e);
    b1 = 13.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 3.0 * (b - d) * (b - d);
    b2 = 13.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) + 3.0 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
  }
  const double eps = 1e-8;
  const double tau = fabs(b0 - b2);
  const double a0 = 1.0 * (1.0 + tau / (eps + b0));
  const double a1 = 6.0 * (1.0 + tau / (eps + b1));
  const double a2 = 3.0 * (1.0 + tau / (eps + b2));
  const double w0 = a0 / (a0 + a1 + a2), w1 = a1 / (a0 + a1 + a2), w2 = a2 / (a0 + a1 + a2);
  return w0 * q0 + w1 * q1 + w2 * q2;
}

// upwinded face flux fh at face P+1/2 along a stride: fm if v[P]+v[P+1] > 0
// else fp (weno_solver.cpp:185-199)
__device__ __forceinline__ double face_flux(const double *__restrict__ f, const double *__restrict__ v,
                                            long c, long st) {
  if (v[c] + v[c + st] > 0.0)
    return weno_face(f[c - 2 * st], f[c - st], f[c], f[c + st], f[c + 2 * st], true);
  return weno_face(f[c - st], f[c], f[c + st], f[c + 2 * st], f[c + 3 * st], false);
}

// derivative along axis a at padded cell c: (fh[P] - fh[P-1]) / h
__device__ __forceinline__ double weno_deriv(const double *__restrict__ f, const double *__restrict__ v, long c,
                                             long st, double h) {
  return (face_flux(f, v, c, st) - face_flux(f, v, c - st, st)) / h;
}

// weno5_find_derivatives on interior cells
template <int DIM>
__global__ void __launch_bounds__(kLsBlock) k_weno_derivs(Grid g, const double *__restrict__ f,
                                                          const double *__restrict__ u, const double *__restrict__ v,
                                                          const double *__restrict__ w, double *fx, double *fy,
                                                          double *fz) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.interior) return;
  const long c = interior_index(g, t);
  fx[c] = weno_deriv(f, u, c, g.s[0], g.h[0]);
  if constexpr (DIM > 1) fy[c] = weno_deriv(f, v, c, g.s[1], g.h[1]);
  if constexpr (DIM > 2) fz[c] = weno_deriv(f, w, c, g.s[2], g.h[2]);
}

// One TVD-RK3 stage fused with the convection RHS (source.cpp:12-39):
// s = -u*fx - v*fy - w*fz on interior cells, then
//   stage 1: out = in + dt*s                 (s1 <- s)
//   stage 2: out = in + dt*(s - 3 s1)/4      (s2 <- s)
//   stage 3: out = in + dt*(-s1 - s2 + 8 s)/12
// fx/fy/fz are also stored, as scalar_data keeps them.
template <int DIM, int STAGE>
__global__ void __launch_bounds__(kLsBlock) k_rk3_stage(Grid g, const double *__restrict__ in, double *__restrict__ out,
                                                        const double *__restrict__ u, const double *__restrict__ v,
                                                        const double *__restrict__ w, double *__restrict__ s1,
                                                        double *__restrict__ s2, double *fx, double *fy, double *fz,
                                                        double dt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.interior) return;
  const long c = interior_index(g, t);
  const double dfx = weno_deriv(in, u, c, g.s[0], g.h[0]);
  double s = -u[c] * dfx;
  fx[c] = dfx;
  if constexpr (DIM > 1) {
    const double dfy = weno_deriv(in, v, c, g.s[1], g.h[1]);
    s -= v[c] * dfy;
    fy[c] = dfy;
  }
  if constexpr (DIM > 2) {
    const double dfz = weno_deriv(in, w, c, g.s[2], g.h[2]);
    s -= w[c] * dfz;
    fz[c] = dfz;
  }
  const double x = in[c];
  if constexpr (STAGE == 1) {
    s1[c] = s;
    out[c] = x + dt * s;
  } else if constexpr (STAGE == 2) {
    s2[c] = s;
    out[c] = x + dt * (s - 3.0 * s1[c]) / 4.0;
  } else {
    out[c] = x + dt * (-s1[c] - s2[c] + 8.0 * s) / 12.0;
  }
}

// ------------------------------------------------------------- BCs
// One boundary line along an axis; p points to padded position 0 of the
// axis, st is the stride.  `face` selects the staggered variant.
__device__ void bc_line(double *p, long st, int n, int N, int type, bool right, bool face, double U, double h) {
  const int g = kGhc;
#define F(q) p[(long)(q) * st]
  if (right) {
    const int I = n - 1 + g;
    switch (type) {
      case SCIHPC_BC_NEUMANN:
        if (face) F(I) = F(I - 1) + U * h;
        for (int i = 1; i <= g; ++i) F(I + i) = F(I + i - 1) + U * h;
        break;
      case SCIHPC_BC_DIRICHLET:
        if (face) F(I) = U;
        else F(I) = F(I - 1) / 3.0 + 2.0 * U / 3.0;
        for (int i = 1; i <= g; ++i) F(I + i) = 2.0 * F(I + i - 1) - F(I + i - 2);
        break;
      case SCIHPC_BC_PERIODIC:
        for (int i = 1; i <= g; ++i) F(I + i) = F(g + (i - 1));
        break;
      case SCIHPC_BC_NO_SLIP:
        if (face) {
          for (int i = 1; i <= g; ++i) F(I + i) = -F(I - i);
          F(I) = 0.0;
        } else {
          for (int i = 1; i <= g; ++i) F(I + i) = -F(I + 1 - i);
        }
        break;
      case SCIHPC_BC_SLIP:
        if (face) {
          for (int i = 1; i <= g; ++i) F(I + i) = F(I - i);
          F(I) = 0.0;
        } else {
          for (int i = 1; i <= g; ++i) F(I + i) = F(I + 1 - i);
        }
        break;
    }
  } else {
    const int I = g;
    switch (type) {
      case SCIHPC_BC_NEUMANN:
        if (face) {
          F(I - 1) = F(I) - U * h;
          for (int i = 1; i <= g - 1; ++i) F(I - 1 - i) = F(I - i) - U * h;
        } else {
          for (int i = 1; i <= g; ++i) F(I - i) = F(I - i + 1) - U * h;
        }
        break;
      case SCIHPC_BC_DIRICHLET:
        if (face) {
          F(I - 1) = U;
          for (int i = 2; i <= g; ++i) F(I - i) = 2.0 * F(I - i + 1) - F(I - i + 2);
        } else {
          F(I) = F(I + 1) / 3.0 + 2.0 * U / 3.0;
          for (int i = 1; i <= g; ++i) F(I - i) = 2.0 * F(I - i + 1) - F(I - i + 2);
        }
        break;
      case SCIHPC_BC_PERIODIC:
        for (int i = 1; i <= g; ++i) F(I - i) = F(N - g - 1 - (i - 1));
        break;
      case SCIHPC_BC_NO_SLIP:
        if (face) {
          for (int i = 1; i <= g; ++i) F(I - i) = -F(I - 2 + i);
          F(I - 1) = 0.0;
        } else {
          for (int i = 1; i <= g; ++i) F(I - i) = -F(I - 1 + i);
        }
        break;
      case SCIHPC_BC_SLIP:
        if (face) {
          for (int i = 1; i <= g; ++i) F(I - i) = F(I - 2 + i);
          F(I - 1) = 0.0;
        } else {
          for (int i = 1; i <= g; ++i) F(I - i) = F(I - 1 + i);
        }
        break;
    }
  }
#undef F
}

struct BcSpec {
  int type[6];
  double value[6];
};

// ghost fill of both sides of axis a, one thread per boundary line
__global__ void __launch_bounds__(kLsBlock) k_bc_axis(Grid g, double *f, BcSpec bc, int a, int stag) {
  const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long lines = (long)g.n[t1] * g.n[t2];
  if (t >= lines) return;
  const int u = (int)(t / g.n[t2]), v = (int)(t % g.n[t2]);
  double *p = f + (long)(u + g.off[t1]) * g.s[t1] + (long)(v + g.off[t2]) * g.s[t2];
  const bool face = stag == a;
  bc_line(p, g.s[a], g.n[a], g.N[a], bc.type[2 * a + 1], true, face, bc.value[2 * a + 1], g.h[a]);
  bc_line(p, g.s[a], g.n[a], g.N[a], bc.type[2 * a], false, face, bc.value[2 * a], g.h[a]);
}

// find_density: rho = H + (1-H) ratio (or mirrored), lsm.cpp:8-18
__global__ void __launch_bounds__(kLsBlock) k_density(const double *__restrict__ phi, long n, double w, double ratio,
                                                      int positive_ref, double *__restrict__ rho) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double x = phi[c];
  const double pi = 3.14159265358979323846;
  double h;
  if (x < -w) h = 0.0;
  else if (x > w) h = 1.0;
  else h = 0.5 * (1 + x / w + sin(pi * x / w) / pi);
  rho[c] = positive_ref ? h + (1.0 - h) * ratio : (1.0 - h) + h * ratio;
}

// ------------------------------------------------- projection step
// The projection kernels round like the reference (no fused multiply-add
// contraction), so their results are bit-identical for the same inputs.
// padded index of interior cell (I, J, K)
__device__ __forceinline__ long pidx(const Grid &g, int I, int J, int K) {
  return (long)(I + g.off[0]) * g.s[0] + (long)(J + g.off[1]) * g.s[1] + (K + g.off[2]);
}
__device__ __forceinline__ void interior_ijk(const Grid &g, long t, int &I, int &J, int &K) {
  // MG order: x fastest
  I = (int)(t % g.n[0]);
  J = (int)((t / g.n[0]) % g.n[1]);
  K = (int)(t / ((long)g.n[0] * g.n[1]));
}

// pressure_guess_solution (projection_method.cpp:803-815), every padded cell
__global__ void __launch_bounds__(kLsBlock) k_pressure_guess(const double *__restrict__ old,
                                                             const double *__restrict__ old2, double *__restrict__ phat,
                                                             long n, double dt_old, double dt_old2) {
#pragma clang fp contract(off)  // the reference rounds every product (no FMA)
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double dpdt = (old[c] - old2[c]) / dt_old2;
  phat[c] = old[c] + dt_old * dpdt;
}

// MG right-hand side of the FPC step (projection_method.cpp:667-692) with
// RHS = sum C p_hat of init_ppe_fpc_coefficients (:467-541) formed in place;
// one thread per interior cell, written in the multigrid layout
template <int DIM>
__global__ void __launch_bounds__(kLsBlock) k_fpc_rhs(Grid g, const double *__restrict__ rho,
                                                      const double *__restrict__ f, const double *__restrict__ u,
                                                      const double *__restrict__ v, const double *__restrict__ w,
                                                      double dt, double rho12, double *__restrict__ mg_rhs) {
#pragma clang fp contract(off)  // the reference rounds every product (no FMA)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.interior) return;
  int I, J, K;
  interior_ijk(g, t, I, J, K);
  const long c = pidx(g, I, J, K), sx = g.s[0], sy = g.s[1], sz = g.s[2];
  const double dx = g.h[0], dy = g.h[1], dz = g.h[2];
  const double R = 0.5 * (rho[c] + rho[c + sx]), L = 0.5 * (rho[c] + rho[c - sx]);
  const double U = 0.5 * (rho[c] + rho[c + sy]), D = 0.5 * (rho[c] + rho[c - sy]);
  double CR = (1.0 - rho12 / R) / dx / dx, CL = (1.0 - rho12 / L) / dx / dx;
  double CU = (1.0 - rho12 / U) / dy / dy, CD = (1.0 - rho12 / D) / dy / dy;
  if (I == 0) CL = 0.0;
  if (I == g.n[0] - 1) CR = 0.0;
  if (J == 0) CD = 0.0;
  if (J == g.n[1] - 1) CU = 0.0;
  double CF = 0.0, CB = 0.0;
  if constexpr (DIM > 2) {
    const double Fz = 0.5 * (rho[c] + rho[c + sz]), B = 0.5 * (rho[c] + rho[c - sz]);
    CF = (1.0 - rho12 / Fz) / dz / dz;
    CB = (1.0 - rho12 / B) / dz / dz;
    if (K == 0) CB = 0.0;
    if (K == g.n[2] - 1) CF = 0.0;
  }
  double CC = -(CR + CL + CU + CD);
  if constexpr (DIM > 2) CC -= (CF + CB);
  double RHS = CC * f[c] + CR * f[c + sx] + CL * f[c - sx] + CU * f[c + sy] + CD * f[c - sy];
  if constexpr (DIM > 2) RHS += CF * f[c + sz] + CB * f[c - sz];
  double r = (u[c] - u[c - sx]) / dx + (v[c] - v[c - sy]) / dy;
  if constexpr (DIM > 2) r += (w[c] - w[c - sz]) / dz;
  r *= rho12 / dt;
  r += RHS;
  mg_rhs[t] = r;
}

// MG right-hand side of the variable-density projection: div(u*) / dt
// (projection_method.cpp:546-568) in the multigrid layout; part[block] = sum
template <int DIM>
__global__ void __launch_bounds__(kLsBlock) k_div_rhs(Grid g, const double *__restrict__ u,
                                                      const double *__restrict__ v, const double *__restrict__ w,
                                                      double dt, double *__restrict__ mg_rhs,
                                                      double *__restrict__ part) {
#pragma clang fp contract(off)  // the reference rounds every product (no FMA)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (t < g.interior) {
    int I, J, K;
    interior_ijk(g, t, I, J, K);
    const long c = pidx(g, I, J, K);
    r = (u[c] - u[c - g.s[0]]) / g.h[0] + (v[c] - v[c - g.s[1]]) / g.h[1];
    if constexpr (DIM > 2) r += (w[c] - w[c - g.s[2]]) / g.h[2];
    r /= dt;
    mg_rhs[t] = r;
  }
  r = block_sum(r);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void __launch_bounds__(kLsBlock) k_sub_const(double *__restrict__ x, long n, double m) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] -= m;
}

// p (padded interior) <- multigrid solution (projection_method.cpp:708-720)
__global__ void __launch_bounds__(kLsBlock) k_mg_to_padded(Grid g, const double *__restrict__ sol,
                                                           double *__restrict__ p) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.interior) return;
  int I, J, K;
  interior_ijk(g, t, I, J, K);
  p[pidx(g, I, J, K)] = sol[t];
}

// velocity corrections over the padded range i < Nx-1, j < Ny-1, k < Nz:
// FPC (projection_method.cpp:723-750) when phat != null, else the
// variable-density projection (projection_method.cpp:630-654)
template <int DIM>
__global__ void __launch_bounds__(kLsBlock) k_correct_velocity(Grid g, const double *__restrict__ rho,
                                                               const double *__restrict__ p,
                                                               const double *__restrict__ ph, double dt,
                                                               double rho12, double *__restrict__ u,
                                                               double *__restrict__ v, double *__restrict__ w) {
#pragma clang fp contract(off)  // the reference rounds every product (no FMA)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ni = g.N[0] - 1, Nj = g.N[1] - 1, Nk = g.N[2];
  if (t >= (long)Ni * Nj * Nk) return;
  const int k = (int)(t % Nk), j = (int)((t / Nk) % Nj), i = (int)(t / ((long)Nk * Nj));
  const long c = (long)i * g.s[0] + (long)j * g.s[1] + k;
  const long sx = g.s[0], sy = g.s[1], sz = g.s[2];
  const double dx = g.h[0], dy = g.h[1], dz = g.h[2];
  if (ph) {
    const double px = (p[c + sx] - p[c]) / dx, px2 = (ph[c + sx] - ph[c]) / dx;
    double density = 0.5 * (rho[c] + rho[c + sx]);
    u[c] -= dt * ((px - px2) / rho12 + px2 / density);
    const double py = (p[c + sy] - p[c]) / dy, py2 = (ph[c + sy] - ph[c]) / dy;
    density = 0.5 * (rho[c] + rho[c + sy]);
    v[c] -= dt * ((py - py2) / rho12 + py2 / density);
    if (DIM > 2 && k < Nk - 1) {
      const double pz = (p[c + sz] - p[c]) / dz, pz2 = (ph[c + sz] - ph[c]) / dz;
      density = 0.5 * (rho[c] + rho[c + sz]);
      w[c] -= dt * ((pz - pz2) / rho12 + pz2 / density);
    }
  } else {
    double density = 0.5 * (rho[c] + rho[c + sx]);
    u[c] -= dt / dx / density * (p[c + sx] - p[c]);
    density = 0.5 * (rho[c] + rho[c + sy]);
    v[c] -= dt / dy / density * (p[c + sy] - p[c]);
    if (DIM > 2 && k < Nk - 1) {
      density = 0.5 * (rho[c] + rho[c + sz]);
      w[c] -= dt / dz / density * (p[c + sz] - p[c]);
    }
  }
}

}  // namespace scihpc

using namespace scihpc;

struct scihpc_ls {
  Grid g{};
  BcSpec bc{};
  double *phi[2] = {nullptr, nullptr};  // ping-pong stage buffers
  int cur = 0;
  double *u = nullptr, *v = nullptr, *w = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *s1 = nullptr, *s2 = nullptr, *tmp = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  double *alloc() {
    void *p = nullptr;
    SCIHPC_HIP(hipMalloc(&p, sizeof(double) * g.total));
    SCIHPC_HIP(hipMemsetAsync(p, 0, sizeof(double) * g.total, stream));
    return static_cast<double *>(p);
  }

  unsigned grid_interior() const { return ceil_div(g.interior, kLsBlock); }

  void apply_bc(double *f, int stag) {
    for (int a = 0; a < g.ndim; ++a) {
      const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
      const long lines = (long)g.n[t1] * g.n[t2];
      k_bc_axis<<<ceil_div(lines, kLsBlock), kLsBlock, 0, stream>>>(g, f, bc, a, stag);
      SCIHPC_CHECK_LAUNCH();
    }
  }

  template <int DIM>
  void rk3_step(double dt) {
    const unsigned nb = grid_interior();
    double *a = phi[cur], *b = phi[cur ^ 1];
    k_rk3_stage<DIM, 1><<<nb, kLsBlock, 0, stream>>>(g, a, b, u, v, w, s1, s2, fx, fy, fz, dt);
    SCIHPC_CHECK_LAUNCH();
    apply_bc(b, -1);
    k_rk3_stage<DIM, 2><<<nb, kLsBlock, 0, stream>>>(g, b, a, u, v, w, s1, s2, fx, fy, fz, dt);
    SCIHPC_CHECK_LAUNCH();
    apply_bc(a, -1);
    k_rk3_stage<DIM, 3><<<nb, kLsBlock, 0, stream>>>(g, a, b, u, v, w, s1, s2, fx, fy, fz, dt);
    SCIHPC_CHECK_LAUNCH();
    apply_bc(b, -1);
    cur ^= 1;
  }

  void step(double dt) {
    if (g.ndim == 1) rk3_step<1>(dt);
    else if (g.ndim == 2) rk3_step<2>(dt);
    else rk3_step<3>(dt);
  }
};

static void ls_check(const scihpc_ls *h) { SCIHPC_REQUIRE(h, "null handle"); }

extern "C" int scihpc_ls_create(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                                const int *bc_types, const double *bc_values, scihpc_ls **out) {
  return guarded([&] {
    SCIHPC_REQUIRE(out, "out is null");
    *out = nullptr;
    SCIHPC_REQUIRE(ndim >= 1 && ndim <= 3, "ndim must be 1, 2 or 3");
    SCIHPC_REQUIRE(nx > 2 && (ndim < 2 || ny > 2) && (ndim < 3 || nz > 2), "each axis needs > 2 cells");
    SCIHPC_REQUIRE(dx > 0 && (ndim < 2 || dy > 0) && (ndim < 3 || dz > 0), "spacings must be positive");
    SCIHPC_REQUIRE(bc_types, "bc_types is null");
    auto *h = new scihpc_ls;
    try {
      h->g = make_grid(ndim, nx, ny, nz, dx, dy, dz);
      for (int q = 0; q < 6; ++q) {
        const int t = bc_types[q];
        SCIHPC_REQUIRE(q >= 2 * ndim || (t >= 0 && t <= 4), "unknown bc type");
        h->bc.type[q] = t;
        h->bc.value[q] = bc_values ? bc_values[q] : 0.0;
      }
      for (int a = 0; a < ndim; ++a)  // bc_factory::check_bc (bc_factory.cpp:16-19)
        SCIHPC_REQUIRE((h->bc.type[2 * a] == SCIHPC_BC_PERIODIC) == (h->bc.type[2 * a + 1] == SCIHPC_BC_PERIODIC),
                       "periodic must be set on both sides of an axis");
      SCIHPC_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
      SCIHPC_HIP(hipEventCreate(&h->ev0));
      SCIHPC_HIP(hipEventCreate(&h->ev1));
      h->phi[0] = h->alloc();
      h->phi[1] = h->alloc();
      h->u = h->alloc();
      h->v = h->alloc();
      h->w = h->alloc();
      h->fx = h->alloc();
      h->fy = h->alloc();
      h->fz = h->alloc();
      h->s1 = h->alloc();
      h->s2 = h->alloc();
      h->tmp = h->alloc();
      SCIHPC_HIP(hipStreamSynchronize(h->stream));
    } catch (...) {
      scihpc_ls_destroy(h);
      throw;
    }
    *out = h;
  });
}

extern "C" int scihpc_ls_destroy(scihpc_ls *h) {
  if (!h) return SCIHPC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (double *p : {h->phi[0], h->phi[1], h->u, h->v, h->w, h->fx, h->fy, h->fz, h->s1, h->s2, h->tmp})
    if (p) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SCIHPC_OK;
}

static void h2d(scihpc_ls *h, double *dst, const double *src) {
  SCIHPC_HIP(hipMemcpyAsync(dst, src, sizeof(double) * h->g.total, hipMemcpyHostToDevice, h->stream));
}
static void d2h(scihpc_ls *h, double *dst, const double *src) {
  SCIHPC_HIP(hipMemcpyAsync(dst, src, sizeof(double) * h->g.total, hipMemcpyDeviceToHost, h->stream));
}

extern "C" int scihpc_ls_set_phi(scihpc_ls *h, const double *phi) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(phi, "phi is null");
    // both stage buffers start identical so ghost corners (never written)
    // keep the caller's values whichever buffer holds the result
    h2d(h, h->phi[0], phi);
    SCIHPC_HIP(hipMemcpyAsync(h->phi[1], h->phi[0], sizeof(double) * h->g.total, hipMemcpyDeviceToDevice,
                              h->stream));
    h->cur = 0;
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_get_phi(scihpc_ls *h, double *phi) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(phi, "phi is null");
    d2h(h, phi, h->phi[h->cur]);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_set_velocity(scihpc_ls *h, const double *u, const double *v, const double *w) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(u && (h->g.ndim < 2 || v) && (h->g.ndim < 3 || w), "velocity component is null");
    h2d(h, h->u, u);
    if (v) h2d(h, h->v, v);
    if (w) h2d(h, h->w, w);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_weno5_derivatives(scihpc_ls *h, const double *flux, double *fx, double *fy, double *fz) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(flux && fx, "null array");
    h2d(h, h->tmp, flux);
    const unsigned nb = h->grid_interior();
    SCIHPC_HIP(hipMemsetAsync(h->fx, 0, sizeof(double) * h->g.total, h->stream));
    SCIHPC_HIP(hipMemsetAsync(h->fy, 0, sizeof(double) * h->g.total, h->stream));
    SCIHPC_HIP(hipMemsetAsync(h->fz, 0, sizeof(double) * h->g.total, h->stream));
    if (h->g.ndim == 1)
      k_weno_derivs<1><<<nb, kLsBlock, 0, h->stream>>>(h->g, h->tmp, h->u, h->v, h->w, h->fx, h->fy, h->fz);
    else if (h->g.ndim == 2)
      k_weno_derivs<2><<<nb, kLsBlock, 0, h->stream>>>(h->g, h->tmp, h->u, h->v, h->w, h->fx, h->fy, h->fz);
    else
      k_weno_derivs<3><<<nb, kLsBlock, 0, h->stream>>>(h->g, h->tmp, h->u, h->v, h->w, h->fx, h->fy, h->fz);
    SCIHPC_CHECK_LAUNCH();
    d2h(h, fx, h->fx);
    if (fy) d2h(h, fy, h->fy);
    if (fz) d2h(h, fz, h->fz);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_tvd_rk3(scihpc_ls *h, double dt, int nsteps) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    for (int s = 0; s < nsteps; ++s) h->step(dt);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_apply_bc(scihpc_ls *h, double *f, int stag) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(f, "null array");
    SCIHPC_REQUIRE(stag >= -1 && stag < h->g.ndim, "stag must be -1 or an axis < ndim");
    h2d(h, h->tmp, f);
    h->apply_bc(h->tmp, stag);
    d2h(h, f, h->tmp);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_density(scihpc_ls *h, double ls_width, double ratio, int positive_ref, double *rho) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(rho && ls_width > 0, "bad density arguments");
    k_density<<<ceil_div(h->g.total, kLsBlock), kLsBlock, 0, h->stream>>>(h->phi[h->cur], h->g.total, ls_width,
                                                                          ratio, positive_ref, h->tmp);
    SCIHPC_CHECK_LAUNCH();
    d2h(h, rho, h->tmp);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_ls_time_step(scihpc_ls *h, double dt, int iters, double *ms, double *bytes) {
  return guarded([&] {
    ls_check(h);
    SCIHPC_REQUIRE(iters > 0, "iters must be > 0");
    h->step(dt);  // warm-up
    SCIHPC_HIP(hipEventRecord(h->ev0, h->stream));
    for (int s = 0; s < iters; ++s) h->step(dt);
    SCIHPC_HIP(hipEventRecord(h->ev1, h->stream));
    SCIHPC_HIP(hipEventSynchronize(h->ev1));
    float t = 0.f;
    SCIHPC_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) *ms = t / iters;
    // minimal Shu-Osher traffic per step: (7 + 3 d) * 8 B per interior cell
    if (bytes) *bytes = (7.0 + 3.0 * h->g.ndim) * 8.0 * (double)h->g.interior;
  });
}

// ================================================= projection step (C ABI)
// One projection step on the padded fields of one grid, around a multigrid
// handle of the same interior grid: the FPC step of SciHPC
// (pressure_guess_solution, init_ppe_fpc_coefficients,
// fast_pressure_correction; projection_method.cpp:467-541, 659-753,
// 803-815) with the reference-faithful full cycles, and the
// variable-density projection (projection_method.cpp:544-657) with the SOR
// replaced by the A(rho) multigrid solve.
struct scihpc_proj {
  Grid g{};
  BcSpec pbc{};     // pressure: Neumann (the pressure wrapper of the drivers)
  BcSpec vbc[3]{};  // velocity components: SLIP -> NO_SLIP wall-normal (wrapper.cpp:111-125)
  hipStream_t stream = nullptr;
  double *rho = nullptr, *pold = nullptr, *pold2 = nullptr, *phat = nullptr, *p = nullptr;
  double *u = nullptr, *v = nullptr, *w = nullptr;
  double *mg_rhs = nullptr, *sol = nullptr, *part = nullptr, *hpart = nullptr;

  double *alloc(long n) {
    void *q = nullptr;
    SCIHPC_HIP(hipMalloc(&q, sizeof(double) * n));
    SCIHPC_HIP(hipMemsetAsync(q, 0, sizeof(double) * n, stream));
    return static_cast<double *>(q);
  }
  void bc(double *f, const BcSpec &b, int stag) {
    for (int a = 0; a < g.ndim; ++a) {
      const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
      const long lines = (long)g.n[t1] * g.n[t2];
      k_bc_axis<<<ceil_div(lines, kLsBlock), kLsBlock, 0, stream>>>(g, f, b, a, stag);
      SCIHPC_CHECK_LAUNCH();
    }
  }
  void h2d(double *dst, const double *src) {
    SCIHPC_HIP(hipMemcpyAsync(dst, src, sizeof(double) * g.total, hipMemcpyHostToDevice, stream));
  }
  void d2h(double *dst, const double *src) {
    SCIHPC_HIP(hipMemcpyAsync(dst, src, sizeof(double) * g.total, hipMemcpyDeviceToHost, stream));
  }
  void check_mg(scihpc_mg *mg) {
    SCIHPC_REQUIRE(mg, "null multigrid handle");
    int a = 0, b = 0, c = 0, deg = 0;
    const int rc = scihpc_mg_level_shape(mg, 0, &a, &b, &c, &deg);
    SCIHPC_REQUIRE(rc == SCIHPC_OK, "bad multigrid handle");
    SCIHPC_REQUIRE(a == g.n[0] && b == g.n[1] && c == (g.ndim > 2 ? g.n[2] : 1), "multigrid grid differs");
  }
  void finish_pressure_and_velocity(scihpc_mg *mg, const double *ph, double dt, double rho12) {
    SCIHPC_REQUIRE(scihpc_mg_get_sol_device(mg, sol) == SCIHPC_OK, scihpc_last_error());
    k_mg_to_padded<<<ceil_div(g.interior, kLsBlock), kLsBlock, 0, stream>>>(g, sol, p);
    SCIHPC_CHECK_LAUNCH();
    bc(p, pbc, -1);
    const long nv = (long)(g.N[0] - 1) * (g.N[1] - 1) * g.N[2];
    if (g.ndim == 2)
      k_correct_velocity<2><<<ceil_div(nv, kLsBlock), kLsBlock, 0, stream>>>(g, rho, p, ph, dt, rho12, u, v, w);
    else
      k_correct_velocity<3><<<ceil_div(nv, kLsBlock), kLsBlock, 0, stream>>>(g, rho, p, ph, dt, rho12, u, v, w);
    SCIHPC_CHECK_LAUNCH();
    bc(u, vbc[0], 0);
    bc(v, vbc[1], 1);
    bc(w, vbc[2], 2);
  }
};

extern "C" int scihpc_proj_create(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                                  const int *vel_bc_types, scihpc_proj **out) {
  return guarded([&] {
    SCIHPC_REQUIRE(out, "out is null");
    *out = nullptr;
    SCIHPC_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
    SCIHPC_REQUIRE(nx > 2 && ny > 2 && (ndim < 3 || nz > 2), "each axis needs > 2 cells");
    SCIHPC_REQUIRE(dx > 0 && dy > 0 && (ndim < 3 || dz > 0), "spacings must be positive");
    SCIHPC_REQUIRE(vel_bc_types, "vel_bc_types is null");
    auto *h = new scihpc_proj;
    try {
      h->g = make_grid(ndim, nx, ny, nz, dx, dy, dz);
      for (int q = 0; q < 6; ++q) {
        const int t = vel_bc_types[q];
        SCIHPC_REQUIRE(q >= 2 * ndim || (t >= 0 && t <= 4 && t != SCIHPC_BC_PERIODIC), "unsupported bc type");
        h->pbc.type[q] = SCIHPC_BC_NEUMANN;
        for (int a = 0; a < 3; ++a) {
          const bool normal = q / 2 == a;
          h->vbc[a].type[q] = (normal && t == SCIHPC_BC_SLIP) ? SCIHPC_BC_NO_SLIP : t;
        }
      }
      SCIHPC_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
      const long T = h->g.total;
      h->rho = h->alloc(T);
      h->pold = h->alloc(T);
      h->pold2 = h->alloc(T);
      h->phat = h->alloc(T);
      h->p = h->alloc(T);
      h->u = h->alloc(T);
      h->v = h->alloc(T);
      h->w = h->alloc(T);
      h->mg_rhs = h->alloc(h->g.interior);
      h->sol = h->alloc(h->g.interior);
      h->part = h->alloc(ceil_div(h->g.interior, kLsBlock));
      SCIHPC_HIP(hipHostMalloc(reinterpret_cast<void **>(&h->hpart),
                               sizeof(double) * ceil_div(h->g.interior, kLsBlock), hipHostMallocDefault));
      SCIHPC_HIP(hipStreamSynchronize(h->stream));
    } catch (...) {
      scihpc_proj_destroy(h);
      throw;
    }
    *out = h;
  });
}

extern "C" int scihpc_proj_destroy(scihpc_proj *h) {
  if (!h) return SCIHPC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (double *q : {h->rho, h->pold, h->pold2, h->phat, h->p, h->u, h->v, h->w, h->mg_rhs, h->sol, h->part})
    if (q) (void)hipFree(q);
  if (h->hpart) (void)hipHostFree(h->hpart);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SCIHPC_OK;
}

extern "C" int scihpc_proj_fpc_step(scihpc_proj *h, scihpc_mg *mg, const scihpc_fpc_params *prm, const double *rho,
                                    const double *p_old, const double *p_old2, double *u, double *v, double *w,
                                    double *p, double *phat, int *cycles) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && prm && rho && p_old && p_old2 && u && v && w && p && phat, "null argument");
    SCIHPC_REQUIRE(prm->dt > 0 && prm->dt_old2 != 0 && prm->rho12 > 0 && prm->max_cycles >= 0, "bad parameters");
    h->check_mg(mg);
    const Grid &g = h->g;
    h->h2d(h->rho, rho);
    h->h2d(h->pold, p_old);
    h->h2d(h->pold2, p_old2);
    h->h2d(h->u, u);
    h->h2d(h->v, v);
    h->h2d(h->w, w);
    k_pressure_guess<<<ceil_div(g.total, kLsBlock), kLsBlock, 0, h->stream>>>(h->pold, h->pold2, h->phat, g.total,
                                                                             prm->dt_old, prm->dt_old2);
    if (g.ndim == 2)
      k_fpc_rhs<2><<<ceil_div(g.interior, kLsBlock), kLsBlock, 0, h->stream>>>(g, h->rho, h->phat, h->u, h->v, h->w,
                                                                             prm->dt, prm->rho12, h->mg_rhs);
    else
      k_fpc_rhs<3><<<ceil_div(g.interior, kLsBlock), kLsBlock, 0, h->stream>>>(g, h->rho, h->phat, h->u, h->v, h->w,
                                                                             prm->dt, prm->rho12, h->mg_rhs);
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    // the loop of fast_pressure_correction (projection_method.cpp:694-706):
    // full cycles from the handle's current solution until the absolute
    // residual is <= ppe_tol
    SCIHPC_REQUIRE(scihpc_mg_set_rhs_device(mg, h->mg_rhs) == SCIHPC_OK, scihpc_last_error());
    double err = 0.0;
    SCIHPC_REQUIRE(scihpc_mg_residual(mg, 0, &err) == SCIHPC_OK, scihpc_last_error());
    int it = 0;
    while (err > prm->ppe_tol && it < prm->max_cycles) {
      ++it;
      SCIHPC_REQUIRE(scihpc_mg_fullcycle_ref(mg) == SCIHPC_OK, scihpc_last_error());
      SCIHPC_REQUIRE(scihpc_mg_residual(mg, 0, &err) == SCIHPC_OK, scihpc_last_error());
    }
    h->finish_pressure_and_velocity(mg, h->phat, prm->dt, prm->rho12);
    h->d2h(u, h->u);
    h->d2h(v, h->v);
    h->d2h(w, h->w);
    h->d2h(p, h->p);
    h->d2h(phat, h->phat);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    if (cycles) *cycles = it;
  });
}

extern "C" int scihpc_proj_step(scihpc_proj *h, scihpc_mg *mg, double dt, const scihpc_solve_opts *opts,
                                const double *rho, double *u, double *v, double *w, double *p,
                                scihpc_solve_stats *stats) {
  int rc = SCIHPC_OK;
  const int g_rc = guarded([&] {
    SCIHPC_REQUIRE(h && opts && rho && u && v && w && p && dt > 0, "null or bad argument");
    h->check_mg(mg);
    const Grid &g = h->g;
    SCIHPC_REQUIRE(scihpc_mg_set_density_padded(mg, rho) == SCIHPC_OK, scihpc_last_error());
    h->h2d(h->rho, rho);
    h->h2d(h->u, u);
    h->h2d(h->v, v);
    h->h2d(h->w, w);
    const unsigned nb = ceil_div(g.interior, kLsBlock);
    if (g.ndim == 2)
      k_div_rhs<2><<<nb, kLsBlock, 0, h->stream>>>(g, h->u, h->v, h->w, dt, h->mg_rhs, h->part);
    else
      k_div_rhs<3><<<nb, kLsBlock, 0, h->stream>>>(g, h->u, h->v, h->w, dt, h->mg_rhs, h->part);
    SCIHPC_CHECK_LAUNCH();
    // compatibility of the all-Neumann problem: remove the mean of div u / dt
    SCIHPC_HIP(hipMemcpyAsync(h->hpart, h->part, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    double sum = 0.0;
    for (unsigned b = 0; b < nb; ++b) sum += h->hpart[b];
    k_sub_const<<<nb, kLsBlock, 0, h->stream>>>(h->mg_rhs, g.interior, sum / (double)g.interior);
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    SCIHPC_REQUIRE(scihpc_mg_set_rhs_device(mg, h->mg_rhs) == SCIHPC_OK, scihpc_last_error());
    scihpc_solve_stats st{};
    rc = scihpc_mg_solve(mg, opts, &st);
    SCIHPC_REQUIRE(rc == SCIHPC_OK || rc == SCIHPC_NOT_CONVERGED, scihpc_last_error());
    h->finish_pressure_and_velocity(mg, nullptr, dt, 1.0);
    h->d2h(u, h->u);
    h->d2h(v, h->v);
    h->d2h(w, h->w);
    h->d2h(p, h->p);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    if (stats) *stats = st;
  });
  return g_rc != SCIHPC_OK ? g_rc : rc;
}
