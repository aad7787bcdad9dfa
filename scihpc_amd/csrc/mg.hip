// mg.hip -- pressure-Poisson multigrid for gfx950 behind the C ABI of
// include/scihpc_hip.h.
//
// Replaces multigrid / multigrid_base (multigrid.cpp, multigrid_base.cpp)
// and the variable-coefficient operator of init_ppe_proj_coefficients
// (projection_method.cpp:404-465).
//
// Storage: every level keeps its vectors unpadded, x fastest (the
// reference's of() layout), so callers' rhs/sol copy in and out unchanged.
// Operator: constant-coefficient levels read no coefficient array; variable
// levels keep one face-coefficient array per axis (cx[c] couples cell c to
// c+1 and is 0 on a wall face), the diagonal is formed on the fly.
//
// Two algorithms share the storage:
//  * the MI355X solver: red-black Gauss-Seidel V(nu1,nu2), fused residual +
//    child-average restriction, cell-centred (bi/tri)linear prolongation,
//    dense coarsest solve (explicit inverse, one workgroup), in FP64 or as
//    an FP32 V-cycle inside an FP64 defect-correction loop;
//  * the reference-faithful pieces (steepest-descent relax, slope
//    prolongation, refinement coarse solve) for drivers that call them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "learned.hpp"

namespace scihpc {

constexpr int kBlock = 256;
constexpr int kMaxPartials = 4096;
constexpr int kMaxCoarse = 2048;  // dense coarsest solve, 2 x 16 KiB of LDS

// ------------------------------------------------------------------ views
template <typename T>
struct DLevel {
  int nx, ny, nz;
  long n;
  T *sol, *rhs, *res;
  const T *cx, *cy, *cz;  // face coefficients (variable mode) or nullptr
  T ax, ay, az;           // constant-mode coefficients 1/dx^2 ...
  T dconst;               // constant Dirichlet diagonal
  int dirichlet;          // constant mode: 1 = init_DirichletBC operator
};

// scalars kept on the device by the reference-faithful path
enum { S_R2 = 0, S_RAR = 1, S_SUM = 2, S_FLAG = 3, S_NORM = 4, S_B2 = 5, S_COUNT = 8 };

// -------------------------------------------------------- stencil helpers
// off = sum over present neighbours of coef*x, diag = A(c,c).  Loads of
// neighbours are guarded (no padding).
template <typename T, int DIM, bool VAR>
__device__ __forceinline__ void stencil(const DLevel<T> &L, const T *__restrict__ x, int i,
                                        int j, int k, long c, T &off, T &diag) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const long sy = nx, sz = (long)nx * ny;
  if constexpr (VAR) {
    const T cR = L.cx[c], cL = i > 0 ? L.cx[c - 1] : T(0);
    const T cU = L.cy[c], cD = j > 0 ? L.cy[c - sy] : T(0);
    T o = T(0);
    if (i > 0) o += cL * x[c - 1];
    if (i < nx - 1) o += cR * x[c + 1];
    if (j > 0) o += cD * x[c - sy];
    if (j < ny - 1) o += cU * x[c + sy];
    T d = -(cR + cL + cU + cD);
    if constexpr (DIM == 3) {
      const T cF = L.cz[c], cB = k > 0 ? L.cz[c - sz] : T(0);
      if (k > 0) o += cB * x[c - sz];
      if (k < nz - 1) o += cF * x[c + sz];
      d -= (cF + cB);
    }
    off = o;
    diag = d;
  } else {
    T o = T(0), cc = T(0);
    if (i > 0) { o += L.ax * x[c - 1]; cc += L.ax; }
    if (i < nx - 1) { o += L.ax * x[c + 1]; cc += L.ax; }
    if (j > 0) { o += L.ay * x[c - sy]; cc += L.ay; }
    if (j < ny - 1) { o += L.ay * x[c + sy]; cc += L.ay; }
    if constexpr (DIM == 3) {
      if (k > 0) { o += L.az * x[c - sz]; cc += L.az; }
      if (k < nz - 1) { o += L.az * x[c + sz]; cc += L.az; }
    }
    off = o;
    diag = L.dirichlet ? L.dconst : -cc;
  }
}

__device__ __forceinline__ void cell_ijk(long c, int nx, int ny, int &i, int &j, int &k) {
  i = (int)(c % nx);
  long r = c / nx;
  j = (int)(r % ny);
  k = (int)(r / ny);
}

// ---------------------------------------------------- red-black smoother
// One colour of Gauss-Seidel: x_c = (b_c - off) / diag for cells with
// (i+j+k)%2 == color.  zero_guess: neighbours read as 0 (first half-sweep
// from a zero initial guess, saves clearing x).
template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_gs_color(DLevel<T> L, int color, int zero_guess) {
  const int hx = (L.nx + 1) >> 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)hx * L.ny * L.nz;
  if (t >= total) return;
  const int i2 = (int)(t % hx);
  const long r = t / hx;
  const int j = (int)(r % L.ny), k = (int)(r / L.ny);
  const int i = 2 * i2 + ((j + k + color) & 1);
  if (i >= L.nx) return;
  const long c = i + (long)L.nx * (j + (long)L.ny * k);
  T off, diag;
  stencil<T, DIM, VAR>(L, L.sol, i, j, k, c, off, diag);
  if (zero_guess) off = T(0);
  L.sol[c] = (L.rhs[c] - off) / diag;
}

// ------------------------------------- fused residual + child average
// coarse.rhs = mean over the r^d children of (b - A x), coarse.sol = 0.
template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_residual_restrict(DLevel<T> F, DLevel<T> G, int r) {
  const long C = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= G.n) return;
  int I, J, K;
  cell_ijk(C, G.nx, G.ny, I, J, K);
  const int rz = DIM == 3 ? r : 1;
  T s = T(0);
  for (int kk = 0; kk < rz; ++kk)
    for (int jj = 0; jj < r; ++jj)
      for (int ii = 0; ii < r; ++ii) {
        const int i = I * r + ii, j = J * r + jj, k = K * rz + kk;
        const long c = i + (long)F.nx * (j + (long)F.ny * k);
        T off, diag;
        stencil<T, DIM, VAR>(F, F.sol, i, j, k, c, off, diag);
        s += F.rhs[c] - (off + diag * F.sol[c]);
      }
  G.rhs[C] = s / T(r * r * rz);
  G.sol[C] = T(0);
}

// ------------------------------------------------ linear prolongation
// Cell-centred linear interpolation per axis: X = (2i+1-r)/(2r), exact
// integer weights; ghost rule Neumann -> nearest interior, Dirichlet -> 0.
__device__ __forceinline__ void interp_axis(int i, int r, int &I0, float &tf, double &td) {
  const int num = 2 * i + 1 - r, den = 2 * r;
  const int q = num >= 0 ? num / den : -((-num + den - 1) / den);
  I0 = q;
  td = double(num - q * den) / double(den);
  tf = float(td);
}

template <typename T>
__device__ __forceinline__ T coarse_at(const DLevel<T> &G, int dir, int I, int J, int K) {
  const bool out = I < 0 || I >= G.nx || J < 0 || J >= G.ny || K < 0 || K >= G.nz;
  if (out && dir) return T(0);
  I = min(max(I, 0), G.nx - 1);
  J = min(max(J, 0), G.ny - 1);
  K = min(max(K, 0), G.nz - 1);
  return G.sol[I + (long)G.nx * (J + (long)G.ny * K)];
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock) k_prolong_add(DLevel<T> G, DLevel<T> F, int r, int dir) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F.n) return;
  int i, j, k;
  cell_ijk(c, F.nx, F.ny, i, j, k);
  int I0, J0, K0 = 0;
  float fx, fy, fz;
  double dx_, dy_, dz_ = 0.0;
  interp_axis(i, r, I0, fx, dx_);
  interp_axis(j, r, J0, fy, dy_);
  const T tx = sizeof(T) == 8 ? T(dx_) : T(fx), ty = sizeof(T) == 8 ? T(dy_) : T(fy);
  T v00 = (T(1) - tx) * coarse_at(G, dir, I0, J0, K0) + tx * coarse_at(G, dir, I0 + 1, J0, K0);
  T v10 = (T(1) - tx) * coarse_at(G, dir, I0, J0 + 1, K0) + tx * coarse_at(G, dir, I0 + 1, J0 + 1, K0);
  T v = (T(1) - ty) * v00 + ty * v10;
  if constexpr (DIM == 3) {
    interp_axis(k, r, K0, fz, dz_);
    const T tz = sizeof(T) == 8 ? T(dz_) : T(fz);
    v00 = (T(1) - tx) * coarse_at(G, dir, I0, J0, K0) + tx * coarse_at(G, dir, I0 + 1, J0, K0);
    v10 = (T(1) - tx) * coarse_at(G, dir, I0, J0 + 1, K0) + tx * coarse_at(G, dir, I0 + 1, J0 + 1, K0);
    const T v0 = (T(1) - ty) * v00 + ty * v10;
    const T v01 = (T(1) - tx) * coarse_at(G, dir, I0, J0, K0 + 1) + tx * coarse_at(G, dir, I0 + 1, J0, K0 + 1);
    const T v11 = (T(1) - tx) * coarse_at(G, dir, I0, J0 + 1, K0 + 1) +
                  tx * coarse_at(G, dir, I0 + 1, J0 + 1, K0 + 1);
    const T v1 = (T(1) - ty) * v01 + ty * v11;
    v = (T(1) - tz) * v0 + tz * v1;
  }
  F.sol[c] += v;
}

// --------------------------------------------- dense coarsest solve
// x = Minv b in FP64 on one workgroup; Neumann: subtract the mean (the
// pinned inverse maps a nonzero-mean b to x0 - mean(b)/s, see set-up).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_coarse_solve(DLevel<T> G, const double *__restrict__ Minv,
                                                         int neumann) {
  __shared__ double xs[kMaxCoarse];
  __shared__ double bs[kMaxCoarse];
  const int n = (int)G.n;
  for (int c = threadIdx.x; c < n; c += blockDim.x) bs[c] = double(G.rhs[c]);
  __syncthreads();
  double local = 0.0;
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    const double *m = Minv + (long)row * n;
    double acc = 0.0;
    for (int c = 0; c < n; ++c) acc += m[c] * bs[c];
    xs[row] = acc;
    local += acc;
  }
  double tot = block_sum(local);
  __shared__ double mean;
  if (threadIdx.x == 0) mean = neumann ? tot / n : 0.0;
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) G.sol[c] = T(xs[c] - mean);
}

// --------------------------------------------------- residual + norm
// res = b - A x (stored when store != 0); partial[block] = sum res^2.
template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_residual(DLevel<T> L, int store, double *__restrict__ partial,
                                                     const double *__restrict__ flag) {
  if (flag && *flag != 0.0) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < L.n) {
    int i, j, k;
    cell_ijk(c, L.nx, L.ny, i, j, k);
    T off, diag;
    stencil<T, DIM, VAR>(L, L.sol, i, j, k, c, off, diag);
    const T r = L.rhs[c] - (off + diag * L.sol[c]);
    if (store) L.res[c] = r;
    v = double(r) * double(r);
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// y = A x; partial[block] = sum x*y
template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_apply_dot(DLevel<T> L, const T *__restrict__ x, T *__restrict__ y,
                                                      double *__restrict__ partial,
                                                      const double *__restrict__ flag) {
  if (flag && *flag != 0.0) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < L.n) {
    int i, j, k;
    cell_ijk(c, L.nx, L.ny, i, j, k);
    T off, diag;
    stencil<T, DIM, VAR>(L, x, i, j, k, c, off, diag);
    const T yc = off + diag * x[c];
    y[c] = yc;
    v = double(x[c]) * double(yc);
  }
  v = block_sum(v);
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = v;
}

// out[slot] = sum of partial[0..nb); op 1: also test the norm against
// 1e-16 and raise the early-exit flag (relax, multigrid_base.cpp:23-25)
__global__ void k_finalize(const double *__restrict__ partial, int nb, double *__restrict__ scal,
                           int slot, int op) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[b];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    if (op == 1) {
      if (scal[S_FLAG] != 0.0) return;
      scal[S_NORM] = sqrt(s);
      if (sqrt(s) < 1e-16) scal[S_FLAG] = 1.0;
    }
    scal[slot] = s;
  }
}

// steepest-descent update sol += alpha res with alpha = rdot / (r^T A r);
// partial[block] = sum of the updated sol (for the compatibility step)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sd_update(DLevel<T> L, const double *__restrict__ scal,
                                                      double *__restrict__ partial) {
  if (scal[S_FLAG] != 0.0) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < L.n) {
    const double alpha = scal[S_R2] / scal[S_RAR];
    L.sol[c] += alpha * L.res[c];
    v = L.sol[c];
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kBlock) k_sum(const double *__restrict__ x, long n, double *__restrict__ partial) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = c < n ? x[c] : 0.0;
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kBlock) k_sumsq(const double *__restrict__ x, long n, double *__restrict__ partial) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = c < n ? x[c] * x[c] : 0.0;
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// x -= scal[slot] / n  (compatibility_condition, multigrid_base.cpp:62-66)
__global__ void __launch_bounds__(kBlock) k_sub_mean(double *__restrict__ x, long n, const double *__restrict__ scal,
                                                     int slot, const double *__restrict__ flag) {
  if (flag && *flag != 0.0) return;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    const double m = scal[slot] / n;
    x[c] -= m;
  }
}

// child average without relaxation (restriction body, multigrid_base.cpp:78-111)
__global__ void __launch_bounds__(kBlock) k_child_avg(DLevel<double> F, DLevel<double> G, int r, int ndim) {
  const long C = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= G.n) return;
  int I, J, K;
  cell_ijk(C, G.nx, G.ny, I, J, K);
  double s = 0.0;
  if (ndim == 2) {
    for (int ii = 0; ii < r; ++ii)
      for (int jj = 0; jj < r; ++jj) s += F.res[(I * r + ii) + (long)F.nx * (J * r + jj)];
    s /= r * r;
  } else {
    for (int ii = 0; ii < r; ++ii)
      for (int jj = 0; jj < r; ++jj)
        for (int kk = 0; kk < r; ++kk)
          s += F.res[(I * r + ii) + (long)F.nx * ((J * r + jj) + (long)F.ny * (K * r + kk))];
    s /= r * r * r;
  }
  G.rhs[C] = s;
  G.sol[C] = 0.0;
}

// slope prolongation (multigrid_base.cpp:115-206), one thread per fine cell
__global__ void __launch_bounds__(kBlock) k_slope_prolong(DLevel<double> G, DLevel<double> F, int r, int ndim,
                                                          double cdx, double cdy, double cdz, double fdx,
                                                          double fdy, double fdz) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F.n) return;
  int i, j, k;
  cell_ijk(c, F.nx, F.ny, i, j, k);
  const int I = i / r, J = j / r, K = ndim == 3 ? k / r : 0;
  const int ii = i - I * r, jj = j - J * r, kk = k - K * r;
  const long sy = G.nx, sz = (long)G.nx * G.ny;
  const long o = I + sy * J + sz * K;
  const double *s = G.sol;
  double gx, gy, gz = 0.0;
  if (I > 0 && I < G.nx - 1) gx = 0.5 * (s[o + 1] - s[o - 1]) / cdx;
  else if (I == 0) gx = (s[o + 1] - s[o]) / cdx;
  else gx = (s[o] - s[o - 1]) / cdx;
  if (J > 0 && J < G.ny - 1) gy = 0.5 * (s[o + sy] - s[o - sy]) / cdy;
  else if (J == 0) gy = (s[o + sy] - s[o]) / cdy;
  else gy = (s[o] - s[o - sy]) / cdy;
  const double x = (0.5 + ii) * fdx, y = (0.5 + jj) * fdy;
  double v;
  if (ndim == 3) {
    if (K > 0 && K < G.nz - 1) gz = 0.5 * (s[o + sz] - s[o - sz]) / cdz;
    else if (K == 0) gz = (s[o + sz] - s[o]) / cdz;
    else gz = (s[o] - s[o - sz]) / cdz;
    const double z = (0.5 + kk) * fdz;
    v = s[o] + gx * (x - 0.5 * cdx) + gy * (y - 0.5 * cdy) + gz * (z - 0.5 * cdz);
  } else {
    v = s[o] + gx * (x - 0.5 * cdx) + gy * (y - 0.5 * cdy);
  }
  F.sol[c] += v;
}

// sol += Minv * res (refinement step of solve(), multigrid_base.cpp:229)
__global__ void __launch_bounds__(kBlock) k_coarse_refine(DLevel<double> G, const double *__restrict__ Minv,
                                                          int add) {
  __shared__ double bs[kMaxCoarse];
  const int n = (int)G.n;
  const double *b = add ? G.res : G.rhs;
  for (int c = threadIdx.x; c < n; c += blockDim.x) bs[c] = b[c];
  __syncthreads();
  for (int row = threadIdx.x; row < n; row += blockDim.x) {
    const double *m = Minv + (long)row * n;
    double acc = 0.0;
    for (int c = 0; c < n; ++c) acc += m[c] * bs[c];
    G.sol[row] = add ? G.sol[row] + acc : acc;
  }
}

// ----------------------------------------- mixed-precision outer step
// p_new = p_old + e;  r = b - A p_new (FP64);  rhs32 = float(r);
// partial[block] = sum r^2.  first != 0: p_old = 0 and e = 0 (r = b).
template <int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_outer_update(DLevel<double> L, const double *__restrict__ p_old,
                                                         const float *__restrict__ e, double *__restrict__ p_new,
                                                         float *__restrict__ r32, double *__restrict__ partial,
                                                         int first) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < L.n) {
    double r;
    if (first) {
      r = L.rhs[c];
      p_new[c] = 0.0;
    } else {
      int i, j, k;
      cell_ijk(c, L.nx, L.ny, i, j, k);
      const long sy = L.nx, sz = (long)L.nx * L.ny;
      auto P = [&](long q) { return p_old[q] + double(e[q]); };
      double off = 0.0, diag;
      if constexpr (VAR) {
        const double cR = L.cx[c], cL = i > 0 ? L.cx[c - 1] : 0.0;
        const double cU = L.cy[c], cD = j > 0 ? L.cy[c - sy] : 0.0;
        if (i > 0) off += cL * P(c - 1);
        if (i < L.nx - 1) off += cR * P(c + 1);
        if (j > 0) off += cD * P(c - sy);
        if (j < L.ny - 1) off += cU * P(c + sy);
        diag = -(cR + cL + cU + cD);
        if constexpr (DIM == 3) {
          const double cF = L.cz[c], cB = k > 0 ? L.cz[c - sz] : 0.0;
          if (k > 0) off += cB * P(c - sz);
          if (k < L.nz - 1) off += cF * P(c + sz);
          diag -= (cF + cB);
        }
      } else {
        double cc = 0.0;
        if (i > 0) { off += L.ax * P(c - 1); cc += L.ax; }
        if (i < L.nx - 1) { off += L.ax * P(c + 1); cc += L.ax; }
        if (j > 0) { off += L.ay * P(c - sy); cc += L.ay; }
        if (j < L.ny - 1) { off += L.ay * P(c + sy); cc += L.ay; }
        if constexpr (DIM == 3) {
          if (k > 0) { off += L.az * P(c - sz); cc += L.az; }
          if (k < L.nz - 1) { off += L.az * P(c + sz); cc += L.az; }
        }
        diag = L.dirichlet ? L.dconst : -cc;
      }
      const double pc = P(c);
      r = L.rhs[c] - (off + diag * pc);
      p_new[c] = pc;
    }
    r32[c] = float(r);
    v = r * r;
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// ------------------------------------------------- operator set-up
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_faces_from_rho(DLevel<double> L, const double *__restrict__ rho,
                                                           double *cx, double *cy, double *cz, double dx,
                                                           double dy, double dz) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  int i, j, k;
  cell_ijk(c, L.nx, L.ny, i, j, k);
  const long sy = L.nx, sz = (long)L.nx * L.ny;
  // projection_method.cpp:416-424, 443-447: 1.0 / R / dx / dx
  cx[c] = i < L.nx - 1 ? 1.0 / (0.5 * (rho[c] + rho[c + 1])) / dx / dx : 0.0;
  cy[c] = j < L.ny - 1 ? 1.0 / (0.5 * (rho[c] + rho[c + sy])) / dy / dy : 0.0;
  if constexpr (DIM == 3) cz[c] = k < L.nz - 1 ? 1.0 / (0.5 * (rho[c] + rho[c + sz])) / dz / dz : 0.0;
}

// coarse face = mean of the fine faces lying on it / r^2
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_faces_restrict(DLevel<double> F, DLevel<double> G, int r,
                                                           double *cx, double *cy, double *cz) {
  const long C = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= G.n) return;
  int I, J, K;
  cell_ijk(C, G.nx, G.ny, I, J, K);
  const int rz = DIM == 3 ? r : 1;
  const double scale = 1.0 / double(r * r);
  const double cnt = DIM == 3 ? double(r * r) : double(r);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int b = 0; b < rz; b++)
    for (int a = 0; a < r; a++) {
      sx += F.cx[((I + 1) * r - 1) + (long)F.nx * ((J * r + a) + (long)F.ny * (K * rz + b))];
      sy += F.cy[(I * r + a) + (long)F.nx * (((J + 1) * r - 1) + (long)F.ny * (K * rz + b))];
      if constexpr (DIM == 3)
        sz += F.cz[(I * r + a) + (long)F.nx * ((J * r + b) + (long)F.ny * ((K + 1) * r - 1))];
    }
  cx[C] = I < G.nx - 1 ? sx / cnt * scale : 0.0;
  cy[C] = J < G.ny - 1 ? sy / cnt * scale : 0.0;
  if constexpr (DIM == 3) cz[C] = K < G.nz - 1 ? sz / cnt * scale : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_to_float(const double *__restrict__ a, float *__restrict__ b, long n) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) b[c] = float(a[c]);
}

// padded [i][j][k] (k fastest, ghost 3 on axes < ndim) -> MG layout
__global__ void __launch_bounds__(kBlock) k_padded_to_mg(const double *__restrict__ p, double *__restrict__ m,
                                                         int nx, int ny, int nz, int ndim) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (long)nx * ny * nz) return;
  int i, j, k;
  cell_ijk(c, nx, ny, i, j, k);
  const int g = 3;
  const long Ny = ny + 2 * g, Nz = ndim == 3 ? nz + 2 * g : 1;
  const int ko = ndim == 3 ? g : 0;
  m[c] = p[((long)(i + g) * Ny + (j + g)) * Nz + (k + ko)];
}

// ============================================================ host side
template <typename T>
struct Level {
  int nx = 0, ny = 0, nz = 0, deg = 1;
  long n = 0;
  double dx = 1, dy = 1, dz = 1;
  T *sol = nullptr, *rhs = nullptr, *res = nullptr, *buf = nullptr;
  T *cx = nullptr, *cy = nullptr, *cz = nullptr;
};

// Zero-filled device allocation.  The fill is ordered on `stream` (the
// handle's non-blocking stream does not synchronise with the null stream).
template <typename T>
static T *dalloc(long n, hipStream_t stream) {
  void *p = nullptr;
  SCIHPC_HIP(hipMalloc(&p, sizeof(T) * std::max<long>(n, 1)));
  SCIHPC_HIP(hipMemsetAsync(p, 0, sizeof(T) * std::max<long>(n, 1), stream));
  return static_cast<T *>(p);
}

}  // namespace scihpc

using namespace scihpc;

struct scihpc_mg {
  int ndim = 2, bc = 0, L = 0;
  bool variable = false;
  std::vector<Level<double>> d;
  std::vector<Level<float>> f;  // FP32 hierarchy for the mixed path (lazy)
  std::vector<int> nocompat;
  double *Minv = nullptr;       // dense inverse of the (pinned) coarsest operator
  double *partial = nullptr;    // per-block partial sums
  double *scal = nullptr;       // device scalars (S_*)
  double *host_scal = nullptr;  // pinned mirror
  double *p_alt = nullptr;      // FP64 ping-pong buffer for the outer loop
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  scihpc::LearnedHook learned;

  // ------------------------------------------------------------ views
  template <typename T>
  DLevel<T> view(const Level<T> &l) const {
    DLevel<T> v;
    v.nx = l.nx;
    v.ny = l.ny;
    v.nz = l.nz;
    v.n = l.n;
    v.sol = l.sol;
    v.rhs = l.rhs;
    v.res = l.res;
    v.cx = variable ? l.cx : nullptr;
    v.cy = variable ? l.cy : nullptr;
    v.cz = variable ? l.cz : nullptr;
    v.ax = T(1.0 / l.dx / l.dx);
    v.ay = T(1.0 / l.dy / l.dy);
    v.az = T(ndim == 3 ? 1.0 / l.dz / l.dz : 0.0);
    v.dirichlet = bc == SCIHPC_MG_DIRICHLET;
    v.dconst = T(ndim == 2 ? -2.0 / l.dx / l.dx - 2.0 / l.dy / l.dy
                           : -2.0 / l.dx / l.dx - 2.0 / l.dy / l.dy - 2.0 / l.dz / l.dz);
    return v;
  }

  // ------------------------------------------------------ dispatchers
#define SCIHPC_DISPATCH(KERNEL, T, GRID, ...)                                          \
  do {                                                                                 \
    if (ndim == 2) {                                                                   \
      if (variable) KERNEL<T, 2, true><<<GRID, kBlock, 0, stream>>>(__VA_ARGS__);      \
      else KERNEL<T, 2, false><<<GRID, kBlock, 0, stream>>>(__VA_ARGS__);              \
    } else {                                                                           \
      if (variable) KERNEL<T, 3, true><<<GRID, kBlock, 0, stream>>>(__VA_ARGS__);      \
      else KERNEL<T, 3, false><<<GRID, kBlock, 0, stream>>>(__VA_ARGS__);              \
    }                                                                                  \
    SCIHPC_CHECK_LAUNCH();                                                             \
  } while (0)

  template <typename T>
  void gs_color(const Level<T> &l, int color, int zero_guess) {
    const long hx = (l.nx + 1) / 2;
    SCIHPC_DISPATCH(k_gs_color, T, ceil_div(hx * l.ny * l.nz, kBlock), view(l), color, zero_guess);
  }
  template <typename T>
  void residual_restrict(const Level<T> &a, const Level<T> &b) {
    SCIHPC_DISPATCH(k_residual_restrict, T, ceil_div(b.n, kBlock), view(a), view(b), b.deg / a.deg);
  }
  template <typename T>
  void prolong_add(const Level<T> &coarse, const Level<T> &fine) {
    const int r = coarse.deg / fine.deg, dir = bc == SCIHPC_MG_DIRICHLET;
    if (ndim == 2)
      k_prolong_add<T, 2><<<ceil_div(fine.n, kBlock), kBlock, 0, stream>>>(view(coarse), view(fine), r, dir);
    else
      k_prolong_add<T, 3><<<ceil_div(fine.n, kBlock), kBlock, 0, stream>>>(view(coarse), view(fine), r, dir);
    SCIHPC_CHECK_LAUNCH();
  }
  template <typename T>
  void coarse_solve(const Level<T> &c) {
    k_coarse_solve<T><<<1, kBlock, 0, stream>>>(view(c), Minv, bc == SCIHPC_MG_NEUMANN);
    SCIHPC_CHECK_LAUNCH();
  }

  // V(nu1,nu2) on hierarchy `lv`, in place on lv[0].sol
  template <typename T>
  void vcycle(std::vector<Level<T>> &lv, int nu1, int nu2, bool zero_guess0) {
    if (L == 1) {
      coarse_solve(lv[0]);
      return;
    }
    for (int l = 0; l < L - 1; ++l) {
      for (int s = 0; s < nu1; ++s) {
        gs_color(lv[l], 0, (l == 0 && zero_guess0 && s == 0) ? 1 : 0);
        gs_color(lv[l], 1, 0);
      }
      if (nu1 == 0 && l == 0 && zero_guess0)
        SCIHPC_HIP(hipMemsetAsync(lv[0].sol, 0, sizeof(T) * lv[0].n, stream));
      residual_restrict(lv[l], lv[l + 1]);
    }
    coarse_solve(lv[L - 1]);
    for (int l = L - 1; l > 0; --l) {
      prolong_add(lv[l], lv[l - 1]);
      for (int s = 0; s < nu2; ++s) {
        gs_color(lv[l - 1], 1, 0);
        gs_color(lv[l - 1], 0, 0);
      }
    }
  }

  // sum of partial[0..nb) -> scal[slot]; returns nothing (device only)
  void finalize(int nb, int slot, int op = 0) {
    k_finalize<<<1, kBlock, 0, stream>>>(partial, nb, scal, slot, op);
    SCIHPC_CHECK_LAUNCH();
  }

  double fetch(int slot) {
    SCIHPC_HIP(hipMemcpyAsync(host_scal, scal, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, stream));
    SCIHPC_HIP(hipStreamSynchronize(stream));
    return host_scal[slot];
  }

  // ||b - A x||_2 on an FP64 level (optionally storing res)
  double residual_norm(int level, bool store) {
    auto &l = d[level];
    const unsigned nb = ceil_div(l.n, kBlock);
    SCIHPC_REQUIRE(nb <= (unsigned)kMaxPartials * 64, "level too large");
    reduce_residual(l, store);
    return std::sqrt(fetch(S_NORM));
  }
  void reduce_residual(const Level<double> &l, bool store, const double *flag = nullptr) {
    const unsigned nb = ceil_div(l.n, kBlock);
    SCIHPC_DISPATCH(k_residual, double, nb, view(l), store ? 1 : 0, partial, flag);
    finalize(nb, S_NORM);
  }

  double vec_norm2(const double *x, long n) {
    const unsigned nb = ceil_div(n, kBlock);
    k_sumsq<<<nb, kBlock, 0, stream>>>(x, n, partial);
    SCIHPC_CHECK_LAUNCH();
    finalize(nb, S_B2);
    return fetch(S_B2);
  }

  void sub_mean(double *x, long n, const double *flag = nullptr) {
    const unsigned nb = ceil_div(n, kBlock);
    k_sum<<<nb, kBlock, 0, stream>>>(x, n, partial);
    SCIHPC_CHECK_LAUNCH();
    finalize(nb, S_SUM);
    k_sub_mean<<<nb, kBlock, 0, stream>>>(x, n, scal, S_SUM, flag);
    SCIHPC_CHECK_LAUNCH();
  }

  // --------------------------------------------- reference-faithful
  void compat(int level, double *x) {
    if (nocompat[level]) return;
    sub_mean(x, d[level].n);
  }

  void relax(int level, int iters) {
    auto &l = d[level];
    const unsigned nb = ceil_div(l.n, kBlock);
    // reset the early-exit flag (scal[S_FLAG] = 0)
    SCIHPC_HIP(hipMemsetAsync(scal + S_FLAG, 0, sizeof(double), stream));
    const double *flag = scal + S_FLAG;
    for (int it = 0; it < iters; ++it) {
      SCIHPC_DISPATCH(k_residual, double, nb, view(l), 1, partial, flag);
      finalize(nb, S_R2, 1);
      SCIHPC_DISPATCH(k_apply_dot, double, nb, view(l), l.res, l.buf, partial, flag);
      finalize(nb, S_RAR);
      k_sd_update<double><<<nb, kBlock, 0, stream>>>(view(l), scal, partial);
      SCIHPC_CHECK_LAUNCH();
      if (!nocompat[level]) {
        finalize(nb, S_SUM);
        k_sub_mean<<<nb, kBlock, 0, stream>>>(l.sol, l.n, scal, S_SUM, flag);
        SCIHPC_CHECK_LAUNCH();
      }
    }
  }

  void restriction_ref(int level) {
    relax(level, 10);
    auto &a = d[level], &b = d[level + 1];
    reduce_residual(a, true);
    k_child_avg<<<ceil_div(b.n, kBlock), kBlock, 0, stream>>>(view(a), view(b), b.deg / a.deg, ndim);
    SCIHPC_CHECK_LAUNCH();
    compat(level + 1, b.rhs);
  }

  void prolongation_ref(int level) {
    auto &c = d[level], &f = d[level - 1];
    k_slope_prolong<<<ceil_div(f.n, kBlock), kBlock, 0, stream>>>(view(c), view(f), c.deg / f.deg, ndim, c.dx,
                                                                  c.dy, c.dz, f.dx, f.dy, f.dz);
    SCIHPC_CHECK_LAUNCH();
    relax(level - 1, 10);
  }

  void coarse_solve_ref() {
    auto &c = d[L - 1];
    k_coarse_refine<<<1, kBlock, 0, stream>>>(view(c), Minv, 0);
    SCIHPC_CHECK_LAUNCH();
    compat(L - 1, c.sol);
    double old_res = -1.0, new_res = 0.0;
    do {
      old_res = old_res < 0.0 ? residual_norm(L - 1, true) : new_res;
      k_coarse_refine<<<1, kBlock, 0, stream>>>(view(c), Minv, 1);
      SCIHPC_CHECK_LAUNCH();
      compat(L - 1, c.sol);
      new_res = residual_norm(L - 1, true);
    } while (1.0 - new_res / old_res > 0.1);
  }

  // ------------------------------------------------ coarse inverse
  void build_coarse_inverse() {
    auto &c = d[L - 1];
    const int n = (int)c.n;
    SCIHPC_REQUIRE(n <= kMaxCoarse, "coarsest level has " + std::to_string(n) +
                                        " unknowns; the dense coarse solve supports <= 2048");
    std::vector<double> cx(n), cy(n), cz(n);
    if (variable) {
      SCIHPC_HIP(hipMemcpyAsync(cx.data(), c.cx, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
      SCIHPC_HIP(hipMemcpyAsync(cy.data(), c.cy, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
      if (ndim == 3)
        SCIHPC_HIP(hipMemcpyAsync(cz.data(), c.cz, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
      SCIHPC_HIP(hipStreamSynchronize(stream));
    }
    const double ax = 1.0 / c.dx / c.dx, ay = 1.0 / c.dy / c.dy, az = ndim == 3 ? 1.0 / c.dz / c.dz : 0.0;
    std::vector<double> A((size_t)n * n, 0.0);
    const long sy = c.nx, sz = (long)c.nx * c.ny;
    for (int kk = 0; kk < c.nz; ++kk)
      for (int j = 0; j < c.ny; ++j)
        for (int i = 0; i < c.nx; ++i) {
          const long q = i + sy * j + sz * kk;
          double *row = &A[(size_t)q * n];
          auto put = [&](long col, double v) { row[col] += v; };
          double diag;
          if (variable) {
            const double cR = cx[q], cL = i > 0 ? cx[q - 1] : 0.0, cU = cy[q], cD = j > 0 ? cy[q - sy] : 0.0;
            if (i > 0) put(q - 1, cL);
            if (i < c.nx - 1) put(q + 1, cR);
            if (j > 0) put(q - sy, cD);
            if (j < c.ny - 1) put(q + sy, cU);
            diag = -(cR + cL + cU + cD);
            if (ndim == 3) {
              const double cF = cz[q], cB = kk > 0 ? cz[q - sz] : 0.0;
              if (kk > 0) put(q - sz, cB);
              if (kk < c.nz - 1) put(q + sz, cF);
              diag -= (cF + cB);
            }
          } else {
            double cc = 0.0;
            if (i > 0) { put(q - 1, ax); cc += ax; }
            if (i < c.nx - 1) { put(q + 1, ax); cc += ax; }
            if (j > 0) { put(q - sy, ay); cc += ay; }
            if (j < c.ny - 1) { put(q + sy, ay); cc += ay; }
            if (ndim == 3) {
              if (kk > 0) { put(q - sz, az); cc += az; }
              if (kk < c.nz - 1) { put(q + sz, az); cc += az; }
            }
            diag = bc == SCIHPC_MG_DIRICHLET ? (ndim == 2 ? -2.0 * ax - 2.0 * ay : -2.0 * ax - 2.0 * ay - 2.0 * az)
                                             : -cc;
          }
          put(q, diag);
        }
    if (bc == SCIHPC_MG_NEUMANN) {
      // pin the constant null space: M = A - (s/n) 1 1^T equals A on
      // mean-free vectors and is nonsingular
      double s = 0.0;
      for (int q = 0; q < n; ++q) s = std::max(s, std::fabs(A[(size_t)q * n + q]));
      for (auto &v : A) v -= s / n;
    }
    // Gauss-Jordan inverse with partial pivoting
    std::vector<double> Inv((size_t)n * n, 0.0);
    for (int q = 0; q < n; ++q) Inv[(size_t)q * n + q] = 1.0;
    for (int col = 0; col < n; ++col) {
      int piv = col;
      for (int r = col + 1; r < n; ++r)
        if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
      SCIHPC_REQUIRE(A[(size_t)piv * n + col] != 0.0, "singular coarsest operator");
      if (piv != col)
        for (int q = 0; q < n; ++q) {
          std::swap(A[(size_t)piv * n + q], A[(size_t)col * n + q]);
          std::swap(Inv[(size_t)piv * n + q], Inv[(size_t)col * n + q]);
        }
      const double inv = 1.0 / A[(size_t)col * n + col];
      for (int q = 0; q < n; ++q) {
        A[(size_t)col * n + q] *= inv;
        Inv[(size_t)col * n + q] *= inv;
      }
      for (int r = 0; r < n; ++r) {
        if (r == col) continue;
        const double fct = A[(size_t)r * n + col];
        if (fct == 0.0) continue;
        for (int q = 0; q < n; ++q) {
          A[(size_t)r * n + q] -= fct * A[(size_t)col * n + q];
          Inv[(size_t)r * n + q] -= fct * Inv[(size_t)col * n + q];
        }
      }
    }
    SCIHPC_HIP(hipMemcpyAsync(Minv, Inv.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, stream));
    SCIHPC_HIP(hipStreamSynchronize(stream));
  }

  // ---------------------------------------------- FP32 hierarchy
  void ensure_fp32() {
    if (!f.empty()) return;
    f.resize(L);
    for (int l = 0; l < L; ++l) {
      auto &a = d[l];
      auto &b = f[l];
      b.nx = a.nx;
      b.ny = a.ny;
      b.nz = a.nz;
      b.deg = a.deg;
      b.n = a.n;
      b.dx = a.dx;
      b.dy = a.dy;
      b.dz = a.dz;
      b.sol = dalloc<float>(b.n, stream);
      b.rhs = dalloc<float>(b.n, stream);
      b.res = dalloc<float>(b.n, stream);
      if (variable) refresh_fp32_faces(l);
    }
  }
  void refresh_fp32_faces(int l) {
    auto &a = d[l];
    auto &b = f[l];
    const unsigned nb = ceil_div(a.n, kBlock);
    if (!b.cx) {
      b.cx = dalloc<float>(b.n, stream);
      b.cy = dalloc<float>(b.n, stream);
      if (ndim == 3) b.cz = dalloc<float>(b.n, stream);
    }
    k_to_float<<<nb, kBlock, 0, stream>>>(a.cx, b.cx, a.n);
    k_to_float<<<nb, kBlock, 0, stream>>>(a.cy, b.cy, a.n);
    if (ndim == 3) k_to_float<<<nb, kBlock, 0, stream>>>(a.cz, b.cz, a.n);
    SCIHPC_CHECK_LAUNCH();
  }

  void set_faces_from_rho(const double *drho) {
    variable = true;
    for (int l = 0; l < L; ++l) {
      auto &a = d[l];
      if (!a.cx) {
        a.cx = dalloc<double>(a.n, stream);
        a.cy = dalloc<double>(a.n, stream);
        if (ndim == 3) a.cz = dalloc<double>(a.n, stream);
      }
    }
    auto &l0 = d[0];
    const unsigned nb = ceil_div(l0.n, kBlock);
    if (ndim == 2)
      k_faces_from_rho<2><<<nb, kBlock, 0, stream>>>(view(l0), drho, l0.cx, l0.cy, l0.cz, l0.dx, l0.dy, l0.dz);
    else
      k_faces_from_rho<3><<<nb, kBlock, 0, stream>>>(view(l0), drho, l0.cx, l0.cy, l0.cz, l0.dx, l0.dy, l0.dz);
    SCIHPC_CHECK_LAUNCH();
    for (int l = 1; l < L; ++l) {
      auto &a = d[l - 1], &b = d[l];
      const unsigned g = ceil_div(b.n, kBlock);
      if (ndim == 2)
        k_faces_restrict<2><<<g, kBlock, 0, stream>>>(view(a), view(b), b.deg / a.deg, b.cx, b.cy, b.cz);
      else
        k_faces_restrict<3><<<g, kBlock, 0, stream>>>(view(a), view(b), b.deg / a.deg, b.cx, b.cy, b.cz);
      SCIHPC_CHECK_LAUNCH();
    }
    SCIHPC_HIP(hipStreamSynchronize(stream));
    if (!f.empty())
      for (int l = 0; l < L; ++l) refresh_fp32_faces(l);
    build_coarse_inverse();
  }

  // ------------------------------------------------------- solve
  int solve(const scihpc_solve_opts &o, scihpc_solve_stats &st) {
    SCIHPC_REQUIRE(o.rel_tol > 0 && o.max_cycles >= 0 && o.nu1 >= 0 && o.nu2 >= 0, "bad solve options");
    auto &l0 = d[0];
    const unsigned nb = ceil_div(l0.n, kBlock);
    SCIHPC_HIP(hipEventRecord(ev0, stream));
    const double bnorm = std::sqrt(vec_norm2(l0.rhs, l0.n));
    const double scale = bnorm > 0.0 ? bnorm : 1.0;
    st = scihpc_solve_stats{};
    int cycles = 0;
    double rr;
    if (o.precision == SCIHPC_PREC_FP64) {
      if (o.zero_guess) SCIHPC_HIP(hipMemsetAsync(l0.sol, 0, sizeof(double) * l0.n, stream));
      if (o.use_learned && learned.active()) learned_initial_guess(st);
      rr = residual_norm(0, false) / scale;
      while (rr > o.rel_tol && cycles < o.max_cycles) {
        vcycle(d, o.nu1, o.nu2, false);
        rr = residual_norm(0, false) / scale;
        ++cycles;
      }
      if (bc == SCIHPC_MG_NEUMANN) sub_mean(l0.sol, l0.n);
    } else {
      ensure_fp32();
      auto &f0 = f[0];
      // p lives in l0.sol / p_alt (ping-pong); e in f0.sol; r32 in f0.rhs
      double *p = l0.sol, *q = p_alt;
      int first = o.zero_guess ? 1 : 0;
      if (!first) SCIHPC_HIP(hipMemsetAsync(f0.sol, 0, sizeof(float) * f0.n, stream));
      if (o.use_learned && learned.active()) {
        if (first) SCIHPC_HIP(hipMemsetAsync(l0.sol, 0, sizeof(double) * l0.n, stream));
        learned_initial_guess(st);
        SCIHPC_HIP(hipMemsetAsync(f0.sol, 0, sizeof(float) * f0.n, stream));
        first = 0;
      }
      auto outer = [&](int is_first) {
        if (ndim == 2) {
          if (variable)
            k_outer_update<2, true><<<nb, kBlock, 0, stream>>>(view(l0), p, f0.sol, q, f0.rhs, partial, is_first);
          else
            k_outer_update<2, false><<<nb, kBlock, 0, stream>>>(view(l0), p, f0.sol, q, f0.rhs, partial, is_first);
        } else {
          if (variable)
            k_outer_update<3, true><<<nb, kBlock, 0, stream>>>(view(l0), p, f0.sol, q, f0.rhs, partial, is_first);
          else
            k_outer_update<3, false><<<nb, kBlock, 0, stream>>>(view(l0), p, f0.sol, q, f0.rhs, partial, is_first);
        }
        SCIHPC_CHECK_LAUNCH();
        finalize(nb, S_NORM);
        std::swap(p, q);
        return std::sqrt(fetch(S_NORM)) / scale;
      };
      rr = outer(first);
      while (rr > o.rel_tol && cycles < o.max_cycles) {
        vcycle(f, o.nu1, o.nu2, true);
        rr = outer(0);
        ++cycles;
      }
      if (p != l0.sol) SCIHPC_HIP(hipMemcpyAsync(l0.sol, p, sizeof(double) * l0.n, hipMemcpyDeviceToDevice, stream));
      if (bc == SCIHPC_MG_NEUMANN) sub_mean(l0.sol, l0.n);
    }
    SCIHPC_HIP(hipEventRecord(ev1, stream));
    SCIHPC_HIP(hipEventSynchronize(ev1));
    float ms = 0.f;
    SCIHPC_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    st.cycles = cycles;
    st.rel_residual = rr;
    st.ms = ms;
    return rr <= o.rel_tol ? SCIHPC_OK : SCIHPC_NOT_CONVERGED;
  }

  // learned initial guess: p = N(b, rho), kept only if it lowers the FP64
  // residual below ||b|| (the zero guess); otherwise p = 0.
  void learned_initial_guess(scihpc_solve_stats &st) {
    auto &l0 = d[0];
    const double b2 = vec_norm2(l0.rhs, l0.n);
    learned.apply(l0.rhs, l0.sol, stream);
    const double r = residual_norm(0, false);
    if (std::isfinite(r) && r * r < b2) {
      st.learned_accepted++;
    } else {
      SCIHPC_HIP(hipMemsetAsync(l0.sol, 0, sizeof(double) * l0.n, stream));
      st.learned_rejected++;
    }
  }
};

// ================================================================ C ABI
namespace scihpc {
static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }
}  // namespace scihpc

extern "C" const char *scihpc_last_error(void) { return scihpc::g_last_error.c_str(); }

static int g_levels(int ndim, int nx, int ny, int nz, std::vector<int> &lx, std::vector<int> &ly,
                    std::vector<int> &lz, std::vector<int> &ld) {
  // multigrid.cpp:8-41: divide by 2, else by 3, while every quotient > 3
  const int min_grid = 3;
  lx = {nx};
  ly = {ny};
  lz = {ndim == 3 ? nz : 1};
  ld = {1};
  while (true) {
    const int cx = lx.back(), cy = ly.back(), cz = lz.back();
    int r = 1;
    for (int cand : {2, 3}) {
      bool ok = cx % cand == 0 && cy % cand == 0 && cx / cand > min_grid && cy / cand > min_grid;
      if (ndim == 3) ok = ok && cz % cand == 0 && cz / cand > min_grid;
      if (ok) {
        r = cand;
        break;
      }
    }
    if (r == 1) break;
    lx.push_back(cx / r);
    ly.push_back(cy / r);
    lz.push_back(ndim == 3 ? cz / r : 1);
    ld.push_back(ld.back() * r);
  }
  return (int)ld.size();
}

extern "C" int scihpc_mg_create(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, int bc,
                                scihpc_mg **out) {
  return guarded([&] {
    SCIHPC_REQUIRE(out, "out is null");
    *out = nullptr;
    SCIHPC_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
    SCIHPC_REQUIRE(nx > 0 && ny > 0 && (ndim == 2 || nz > 0), "grid sizes must be positive");
    SCIHPC_REQUIRE(dx > 0 && dy > 0 && (ndim == 2 || dz > 0), "spacings must be positive");
    SCIHPC_REQUIRE(bc == SCIHPC_MG_NEUMANN || bc == SCIHPC_MG_DIRICHLET, "bc must be NEUMANN or DIRICHLET");
    auto *h = new scihpc_mg;
    try {
      h->ndim = ndim;
      h->bc = bc;
      std::vector<int> lx, ly, lz, ld;
      h->L = g_levels(ndim, nx, ny, nz, lx, ly, lz, ld);
      SCIHPC_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
      SCIHPC_HIP(hipEventCreate(&h->ev0));
      SCIHPC_HIP(hipEventCreate(&h->ev1));
      h->d.resize(h->L);
      h->nocompat.assign(h->L, bc == SCIHPC_MG_DIRICHLET ? 1 : 0);
      for (int l = 0; l < h->L; ++l) {
        auto &v = h->d[l];
        v.nx = lx[l];
        v.ny = ly[l];
        v.nz = lz[l];
        v.deg = ld[l];
        v.n = (long)v.nx * v.ny * v.nz;
        v.dx = dx * ld[l];
        v.dy = dy * ld[l];
        v.dz = ndim == 3 ? dz * ld[l] : 1.0;
        v.sol = dalloc<double>(v.n, h->stream);
        v.rhs = dalloc<double>(v.n, h->stream);
        v.res = dalloc<double>(v.n, h->stream);
        v.buf = dalloc<double>(v.n, h->stream);
      }
      h->Minv = dalloc<double>((long)std::min<long>(h->d.back().n, kMaxCoarse) *
                               std::min<long>(h->d.back().n, kMaxCoarse), h->stream);
      h->partial = dalloc<double>(ceil_div(h->d[0].n, kBlock) + 64, h->stream);
      h->scal = dalloc<double>(S_COUNT, h->stream);
      h->p_alt = dalloc<double>(h->d[0].n, h->stream);
      SCIHPC_HIP(hipHostMalloc((void **)&h->host_scal, sizeof(double) * S_COUNT, hipHostMallocDefault));
      h->build_coarse_inverse();
    } catch (...) {
      scihpc_mg_destroy(h);
      throw;
    }
    *out = h;
  });
}

extern "C" int scihpc_mg_destroy(scihpc_mg *h) {
  if (!h) return SCIHPC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  auto fr = [](void *p) {
    if (p) (void)hipFree(p);
  };
  for (auto &v : h->d) {
    fr(v.sol);
    fr(v.rhs);
    fr(v.res);
    fr(v.buf);
    fr(v.cx);
    fr(v.cy);
    fr(v.cz);
  }
  for (auto &v : h->f) {
    fr(v.sol);
    fr(v.rhs);
    fr(v.res);
    fr(v.cx);
    fr(v.cy);
    fr(v.cz);
  }
  fr(h->Minv);
  fr(h->partial);
  fr(h->scal);
  fr(h->p_alt);
  if (h->host_scal) (void)hipHostFree(h->host_scal);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SCIHPC_OK;
}

extern "C" int scihpc_mg_num_levels(const scihpc_mg *h, int *levels) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && levels, "null argument");
    *levels = h->L;
  });
}

extern "C" int scihpc_mg_level_shape(const scihpc_mg *h, int level, int *nx, int *ny, int *nz, int *degree) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && level >= 0 && level < h->L, "bad level");
    const auto &v = h->d[level];
    if (nx) *nx = v.nx;
    if (ny) *ny = v.ny;
    if (nz) *nz = v.nz;
    if (degree) *degree = v.deg;
  });
}

static void check_level(const scihpc_mg *h, int level) {
  SCIHPC_REQUIRE(h, "null handle");
  SCIHPC_REQUIRE(level >= 0 && level < h->L, "level out of range");
}

extern "C" int scihpc_mg_set_density(scihpc_mg *h, const double *rho) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(rho, "rho is null");
    SCIHPC_REQUIRE(h->bc == SCIHPC_MG_NEUMANN, "A(rho) is defined with wall (Neumann) faces only");
    const long n = h->d[0].n;
    double *drho = dalloc<double>(n, h->stream);
    SCIHPC_HIP(hipMemcpyAsync(drho, rho, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    try {
      h->set_faces_from_rho(drho);
    } catch (...) {
      (void)hipFree(drho);
      throw;
    }
    SCIHPC_HIP(hipFree(drho));
  });
}

extern "C" int scihpc_mg_set_density_padded(scihpc_mg *h, const double *rho_padded) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(rho_padded, "rho is null");
    SCIHPC_REQUIRE(h->bc == SCIHPC_MG_NEUMANN, "A(rho) is defined with wall (Neumann) faces only");
    auto &l0 = h->d[0];
    const int g = 3;
    const long np = (long)(l0.nx + 2 * g) * (l0.ny + 2 * g) * (h->ndim == 3 ? l0.nz + 2 * g : 1);
    double *dp = dalloc<double>(np, h->stream);
    double *drho = dalloc<double>(l0.n, h->stream);
    try {
      SCIHPC_HIP(hipMemcpyAsync(dp, rho_padded, sizeof(double) * np, hipMemcpyHostToDevice, h->stream));
      k_padded_to_mg<<<ceil_div(l0.n, kBlock), kBlock, 0, h->stream>>>(dp, drho, l0.nx, l0.ny, l0.nz, h->ndim);
      SCIHPC_CHECK_LAUNCH();
      h->set_faces_from_rho(drho);
    } catch (...) {
      (void)hipFree(dp);
      (void)hipFree(drho);
      throw;
    }
    SCIHPC_HIP(hipFree(dp));
    SCIHPC_HIP(hipFree(drho));
  });
}

static int copy_vec(scihpc_mg *h, int level, int which, const double *src, double *dst) {
  return guarded([&] {
    check_level(h, level);
    auto &v = h->d[level];
    double *dv = which == 0 ? v.rhs : which == 1 ? v.sol : v.res;
    if (src) {
      SCIHPC_HIP(hipMemcpyAsync(dv, src, sizeof(double) * v.n, hipMemcpyHostToDevice, h->stream));
    } else {
      SCIHPC_REQUIRE(dst, "null destination");
      SCIHPC_HIP(hipMemcpyAsync(dst, dv, sizeof(double) * v.n, hipMemcpyDeviceToHost, h->stream));
    }
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_set_rhs(scihpc_mg *h, int level, const double *src) {
  if (!src) return SCIHPC_BAD_ARG;
  return copy_vec(h, level, 0, src, nullptr);
}
extern "C" int scihpc_mg_get_rhs(scihpc_mg *h, int level, double *dst) { return copy_vec(h, level, 0, nullptr, dst); }
extern "C" int scihpc_mg_set_sol(scihpc_mg *h, int level, const double *src) {
  if (!src) return SCIHPC_BAD_ARG;
  return copy_vec(h, level, 1, src, nullptr);
}
extern "C" int scihpc_mg_get_sol(scihpc_mg *h, int level, double *dst) { return copy_vec(h, level, 1, nullptr, dst); }
extern "C" int scihpc_mg_get_res(scihpc_mg *h, int level, double *dst) { return copy_vec(h, level, 2, nullptr, dst); }

extern "C" int scihpc_mg_set_rhs_device(scihpc_mg *h, const double *dsrc) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(dsrc, "null source");
    SCIHPC_HIP(hipMemcpyAsync(h->d[0].rhs, dsrc, sizeof(double) * h->d[0].n, hipMemcpyDeviceToDevice, h->stream));
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_get_sol_device(scihpc_mg *h, double *ddst) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(ddst, "null destination");
    SCIHPC_HIP(hipMemcpyAsync(ddst, h->d[0].sol, sizeof(double) * h->d[0].n, hipMemcpyDeviceToDevice, h->stream));
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_set_nocompat(scihpc_mg *h, int level, int flag) {
  return guarded([&] {
    check_level(h, level);
    h->nocompat[level] = flag ? 1 : 0;
  });
}

extern "C" int scihpc_mg_residual(scihpc_mg *h, int level, double *norm) {
  return guarded([&] {
    check_level(h, level);
    const double r = h->residual_norm(level, true);
    if (norm) *norm = r;
  });
}

extern "C" int scihpc_mg_apply(scihpc_mg *h, int level, const double *x, double *y) {
  return guarded([&] {
    check_level(h, level);
    SCIHPC_REQUIRE(x && y, "null vector");
    auto &l = h->d[level];
    SCIHPC_HIP(hipMemcpyAsync(l.res, x, sizeof(double) * l.n, hipMemcpyHostToDevice, h->stream));
    const unsigned nb = ceil_div(l.n, kBlock);
    auto &hh = *h;
    const int ndim = hh.ndim;
    const bool variable = hh.variable;
    hipStream_t stream = hh.stream;
    if (ndim == 2) {
      if (variable) k_apply_dot<double, 2, true><<<nb, kBlock, 0, stream>>>(hh.view(l), l.res, l.buf, nullptr, nullptr);
      else k_apply_dot<double, 2, false><<<nb, kBlock, 0, stream>>>(hh.view(l), l.res, l.buf, nullptr, nullptr);
    } else {
      if (variable) k_apply_dot<double, 3, true><<<nb, kBlock, 0, stream>>>(hh.view(l), l.res, l.buf, nullptr, nullptr);
      else k_apply_dot<double, 3, false><<<nb, kBlock, 0, stream>>>(hh.view(l), l.res, l.buf, nullptr, nullptr);
    }
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipMemcpyAsync(y, l.buf, sizeof(double) * l.n, hipMemcpyDeviceToHost, stream));
    SCIHPC_HIP(hipStreamSynchronize(stream));
  });
}

extern "C" int scihpc_mg_compat(scihpc_mg *h, int level, int which) {
  return guarded([&] {
    check_level(h, level);
    SCIHPC_REQUIRE(which == 0 || which == 1, "which must be 0 (rhs) or 1 (sol)");
    h->compat(level, which == 0 ? h->d[level].rhs : h->d[level].sol);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_relax(scihpc_mg *h, int level, int iters) {
  return guarded([&] {
    check_level(h, level);
    SCIHPC_REQUIRE(iters >= 0, "iters must be >= 0");
    h->relax(level, iters);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_restriction(scihpc_mg *h, int level) {
  return guarded([&] {
    check_level(h, level);
    SCIHPC_REQUIRE(level + 1 < h->L, "no coarser level");
    h->restriction_ref(level);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_prolongation(scihpc_mg *h, int level) {
  return guarded([&] {
    check_level(h, level);
    SCIHPC_REQUIRE(level > 0, "no finer level");
    h->prolongation_ref(level);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_coarse_solve(scihpc_mg *h) {
  return guarded([&] {
    check_level(h, 0);
    h->coarse_solve_ref();
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_vcycle_ref(scihpc_mg *h) {
  return guarded([&] {
    check_level(h, 0);
    for (int i = 0; i < h->L - 1; ++i) h->restriction_ref(i);
    h->coarse_solve_ref();
    for (int i = h->L - 1; i > 0; --i) h->prolongation_ref(i);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_fullcycle_ref(scihpc_mg *h) {
  return guarded([&] {
    check_level(h, 0);
    const int L = h->L;
    for (int i = 0; i < L - 1; ++i) h->restriction_ref(i);
    h->coarse_solve_ref();
    for (int iter = 1; iter < L; ++iter) {
      for (int i = L - 1 - iter; i < L - 1; ++i) h->restriction_ref(i);
      h->coarse_solve_ref();
    }
    for (int i = L - 1; i > 0; --i) h->prolongation_ref(i);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_vcycle(scihpc_mg *h, int nu1, int nu2) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(nu1 >= 0 && nu2 >= 0, "sweeps must be >= 0");
    h->vcycle(h->d, nu1, nu2, false);
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_mg_solve(scihpc_mg *h, const scihpc_solve_opts *opts, scihpc_solve_stats *stats) {
  int rc = SCIHPC_OK;
  int g = guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(opts, "null options");
    SCIHPC_REQUIRE(opts->precision == SCIHPC_PREC_FP64 || opts->precision == SCIHPC_PREC_MIXED,
                   "precision must be FP64 or MIXED");
    scihpc_solve_stats st{};
    rc = h->solve(*opts, st);
    if (stats) *stats = st;
  });
  return g != SCIHPC_OK ? g : rc;
}

extern "C" int scihpc_mg_time_kernel(scihpc_mg *h, int which, int iters, double *ms, double *bytes) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(iters > 0 && which >= 0 && which <= 3, "bad timing request");
    auto &l0 = h->d[0];
    const double n = (double)l0.n;
    const int c = h->variable ? 1 : 0;
    const int cf = h->variable ? h->ndim : 0;  // face arrays read per launch
    if (which == 1) h->ensure_fp32();
    auto launch = [&] {
      switch (which) {
        case 0:
          h->gs_color(l0, 0, 0);
          h->gs_color(l0, 1, 0);
          break;
        case 1:
          h->gs_color(h->f[0], 0, 0);
          h->gs_color(h->f[0], 1, 0);
          break;
        case 2: {
          h->ensure_fp32();
          const unsigned nb = ceil_div(l0.n, kBlock);
          auto &f0 = h->f[0];
          if (h->ndim == 2) {
            if (h->variable)
              k_outer_update<2, true><<<nb, kBlock, 0, h->stream>>>(h->view(l0), l0.sol, f0.sol, h->p_alt, f0.rhs,
                                                                   h->partial, 0);
            else
              k_outer_update<2, false><<<nb, kBlock, 0, h->stream>>>(h->view(l0), l0.sol, f0.sol, h->p_alt, f0.rhs,
                                                                    h->partial, 0);
          } else {
            if (h->variable)
              k_outer_update<3, true><<<nb, kBlock, 0, h->stream>>>(h->view(l0), l0.sol, f0.sol, h->p_alt, f0.rhs,
                                                                   h->partial, 0);
            else
              k_outer_update<3, false><<<nb, kBlock, 0, h->stream>>>(h->view(l0), l0.sol, f0.sol, h->p_alt, f0.rhs,
                                                                    h->partial, 0);
          }
          SCIHPC_CHECK_LAUNCH();
          break;
        }
        case 3:
          if (h->L > 1) h->residual_restrict(l0, h->d[1]);
          break;
      }
    };
    // algorithmic bytes per launch (each array read once, written once)
    double b = 0;
    switch (which) {
      case 0: b = (3.0 + cf) * 8.0 * n; break;                 // sol r+w, rhs, faces
      case 1: b = (3.0 + cf) * 4.0 * n; break;
      case 2: b = (8.0 + 4.0 + 8.0 + 8.0 * cf + 8.0 + 4.0) * n; break;  // p,e,b,faces -> p',r32
      case 3: b = (2.0 + cf) * 8.0 * n + 8.0 * n / (h->ndim == 3 ? 8.0 : 4.0) * 2.0; break;
    }
    (void)c;
    launch();  // warm-up
    SCIHPC_HIP(hipEventRecord(h->ev0, h->stream));
    for (int it = 0; it < iters; ++it) launch();
    SCIHPC_HIP(hipEventRecord(h->ev1, h->stream));
    SCIHPC_HIP(hipEventSynchronize(h->ev1));
    float t = 0.f;
    SCIHPC_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) *ms = t / iters;
    if (bytes) *bytes = b;
  });
}
