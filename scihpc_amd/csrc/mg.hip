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
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
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
  // Dirichlet coarse levels of the red-black solver: the reference's stencil
  // zeroes the value at the level-0 ghost centre, so a level of spacing H
  // keeps that boundary with the wall term w = 2/(H^2 (1 + h0/H))
  int rbwall;
  T wx, wy, wz;
};

// scalars kept on the device by the reference-faithful path
enum { S_R2 = 0, S_RAR = 1, S_SUM = 2, S_FLAG = 3, S_NORM = 4, S_B2 = 5, S_ZQ = 8, S_PQ = 9, S_PR = 10, S_COUNT = 16 };

// -------------------------------------------------------- stencil helpers
// off = sum over present neighbours of coef*x, diag = A(c,c).  There is no
// padding: a missing neighbour's load is redirected to the cell itself and
// its coefficient is 0, so every load issues unconditionally (one wait for
// all of them) and adding the 0 * x term leaves the sum unchanged (it starts
// at +0, so no -0 can appear).  The accumulation order is the reference's.
// I: index type (int for the LDS-resident levels of the one-workgroup cycle).
template <typename T, int DIM, bool VAR, typename I = long>
__device__ __forceinline__ void stencil(const DLevel<T> &L, const T *__restrict__ x, int i,
                                        int j, int k, I c, T &off, T &diag) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const I sy = nx, sz = (I)nx * ny;
  const bool hl = i > 0, hr = i < nx - 1, hd = j > 0, hu = j < ny - 1;
  const I cl = hl ? c - 1 : c, cr = hr ? c + 1 : c, cd = hd ? c - sy : c, cu = hu ? c + sy : c;
  const T xl = x[cl], xr = x[cr], xd = x[cd], xu = x[cu];
  if constexpr (VAR) {
    const T cR = L.cx[c], cLv = L.cx[cl], cU = L.cy[c], cDv = L.cy[cd];
    const T cL = hl ? cLv : T(0), cD = hd ? cDv : T(0);
    T o = T(0);
    o += cL * xl;
    o += (hr ? cR : T(0)) * xr;
    o += cD * xd;
    o += (hu ? cU : T(0)) * xu;
    T d = -(cR + cL + cU + cD);
    if constexpr (DIM == 3) {
      const bool hb = k > 0, hf = k < nz - 1;
      const I cb = hb ? c - sz : c, cf = hf ? c + sz : c;
      const T cF = L.cz[c], cBv = L.cz[cb];
      const T cB = hb ? cBv : T(0);
      o += cB * x[cb];
      o += (hf ? cF : T(0)) * x[cf];
      d -= (cF + cB);
    }
    off = o;
    diag = d;
  } else {
    const T wx = L.rbwall ? L.wx : T(0), wy = L.rbwall ? L.wy : T(0);
    T o = T(0), cc = T(0);
    o += (hl ? L.ax : T(0)) * xl;
    cc += hl ? L.ax : wx;
    o += (hr ? L.ax : T(0)) * xr;
    cc += hr ? L.ax : wx;
    o += (hd ? L.ay : T(0)) * xd;
    cc += hd ? L.ay : wy;
    o += (hu ? L.ay : T(0)) * xu;
    cc += hu ? L.ay : wy;
    if constexpr (DIM == 3) {
      const T wz = L.rbwall ? L.wz : T(0);
      const bool hb = k > 0, hf = k < nz - 1;
      const I cb = hb ? c - sz : c, cf = hf ? c + sz : c;
      o += (hb ? L.az : T(0)) * x[cb];
      cc += hb ? L.az : wz;
      o += (hf ? L.az : T(0)) * x[cf];
      cc += hf ? L.az : wz;
    }
    off = o;
    diag = (L.dirichlet && !L.rbwall) ? L.dconst : -cc;
  }
}

// 2-D five-point stencil from explicit neighbour values and coefficients
// (used by the LDS-tiled kernels): off-diagonal sum and diagonal.
template <typename T, bool VAR>
__device__ __forceinline__ void stencil2(const DLevel<T> &L, int gx, int gy, T xl, T xr, T xd, T xu, T cL, T cR,
                                         T cD, T cU, T &off, T &diag) {
  const int nx = L.nx, ny = L.ny;
  T o = T(0);
  if constexpr (VAR) {
    if (gx > 0) o += cL * xl;
    if (gx < nx - 1) o += cR * xr;
    if (gy > 0) o += cD * xd;
    if (gy < ny - 1) o += cU * xu;
    diag = -(cR + cL + cU + cD);
  } else {
    T cc = T(0);
    if (gx > 0) { o += L.ax * xl; cc += L.ax; } else if (L.rbwall) cc += L.wx;
    if (gx < nx - 1) { o += L.ax * xr; cc += L.ax; } else if (L.rbwall) cc += L.wx;
    if (gy > 0) { o += L.ay * xd; cc += L.ay; } else if (L.rbwall) cc += L.wy;
    if (gy < ny - 1) { o += L.ay * xu; cc += L.ay; } else if (L.rbwall) cc += L.wy;
    diag = (L.dirichlet && !L.rbwall) ? L.dconst : -cc;
  }
  off = o;
}

// 32-bit index arithmetic: every level has < 2^31 cells, and 64-bit
// division has no hardware path on CDNA
__device__ __forceinline__ void cell_ijk(long c, int nx, int ny, int &i, int &j, int &k) {
  const unsigned u = (unsigned)c, ux = (unsigned)nx, uy = (unsigned)ny;
  const unsigned r = u / ux;
  i = (int)(u - r * ux);
  j = (int)(r % uy);
  k = (int)(r / uy);
}

// 1 / s for the smoothers and transfers.  FP32: the hardware reciprocal
// (v_rcp_f32, about 1 ulp; the FP32 V-cycle only preconditions the FP64
// iteration).  FP64: the hardware estimate refined by two Newton-Raphson
// steps r <- r + r (1 - s r), which square the relative error each time and
// end within an ulp or two of 1/s, in 5 instructions instead of the ~10 of
// an IEEE division.  s = +inf gives 0 in both.
template <typename T>
__device__ __forceinline__ T smooth_rcp(T s) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_amdgcn_rcpf(s);
  } else {
    const double r0 = __builtin_amdgcn_rcp(s);
    double e = __builtin_fma(-s, r0, 1.0);
    double r = __builtin_fma(r0, e, r0);
    e = __builtin_fma(-s, r, 1.0);
    r = __builtin_fma(r, e, r);
    return r0 == 0.0 ? 0.0 : r;
  }
}

// a / d for a Gauss-Seidel update.  FP32: a times the hardware reciprocal.
// FP64: a times the refined reciprocal, then one residual correction
// q + r (a - d q) (Markstein's quotient step), which rounds like the IEEE
// division in nearly every case; the smoothers' iterates agree with the
// oracle's to rounding, as the tests bound them.
template <typename T>
__device__ __forceinline__ T gs_div(T a, T d) {
  if constexpr (sizeof(T) == 4) {
    return a * __builtin_amdgcn_rcpf(d);
  } else {
    const double r = smooth_rcp<double>(d);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-d, q, a), r, q);
  }
}

// ---------------------------------------------------- red-black smoother
// Gauss-Seidel update of cell (i,j,k): x_c = (b_c - off) / diag.
// zero_guess: neighbours read as 0 (first half-sweep from a zero initial
// guess, which saves clearing x).
template <typename T, int DIM, bool VAR, typename I = long>
__device__ __forceinline__ void gs_cell(const DLevel<T> &L, int i, int j, int k, int zero_guess) {
  const I c = i + (I)L.nx * (j + (I)L.ny * k);
  T off, diag;
  stencil<T, DIM, VAR, I>(L, L.sol, i, j, k, c, off, diag);
  if (zero_guess) off = T(0);
  L.sol[c] = gs_div<T>(L.rhs[c] - off, diag);
}

// t-th cell of colour `color` ((i+j+k)%2 == color); false when out of range
__device__ __forceinline__ bool color_cell(long t, int nx, int ny, int nz, int color, int &i, int &j, int &k) {
  const unsigned hx = (unsigned)((nx + 1) >> 1);
  const unsigned u = (unsigned)t;
  if (t >= (long)hx * ny * nz) return false;
  const unsigned r = u / hx;
  const int i2 = (int)(u - r * hx);
  j = (int)(r % (unsigned)ny);
  k = (int)(r / (unsigned)ny);
  i = 2 * i2 + ((j + k + color) & 1);
  return i < nx;
}

template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_gs_color(DLevel<T> L, int color, int zero_guess) {
  int i, j, k;
  if (color_cell((long)blockIdx.x * blockDim.x + threadIdx.x, L.nx, L.ny, L.nz, color, i, j, k))
    gs_cell<T, DIM, VAR>(L, i, j, k, zero_guess);
}

// ------------------------------------- fused residual + child average
// coarse.rhs = mean over the r^d children of (b - A x), coarse.sol = 0.
template <typename T, int DIM, bool VAR, typename Ix = long>
__device__ __forceinline__ void rr_cell(const DLevel<T> &F, const DLevel<T> &G, int r, Ix C) {
  int I, J, K;
  cell_ijk(C, G.nx, G.ny, I, J, K);
  const int rz = DIM == 3 ? r : 1;
  T s = T(0);
  for (int kk = 0; kk < rz; ++kk)
    for (int jj = 0; jj < r; ++jj)
      for (int ii = 0; ii < r; ++ii) {
        const int i = I * r + ii, j = J * r + jj, k = K * rz + kk;
        const Ix c = i + (Ix)F.nx * (j + (Ix)F.ny * k);
        T off, diag;
        stencil<T, DIM, VAR, Ix>(F, F.sol, i, j, k, c, off, diag);
        s += F.rhs[c] - (off + diag * F.sol[c]);
      }
  G.rhs[C] = s / T(r * r * rz);
  G.sol[C] = T(0);
}

template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_residual_restrict(DLevel<T> F, DLevel<T> G, int r) {
  const long C = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (C < G.n) rr_cell<T, DIM, VAR>(F, G, r, C);
}

// ------------------------------------------------ linear prolongation
// Cell-centred linear interpolation per axis: X = (2i+1-r)/(2r), exact
// integer weights; ghost rule Neumann -> nearest interior, Dirichlet -> 0.
__device__ __forceinline__ void interp_axis(int i, int r, int &I0, float &tf, double &td) {
  if (r == 2) {  // num = 2i - 1, den = 4: the same values without the divisions
    I0 = (i - 1) >> 1;
    td = (i & 1) ? 0.25 : 0.75;
    tf = float(td);
    return;
  }
  const int num = 2 * i + 1 - r, den = 2 * r;
  const int q = num >= 0 ? num / den : -((-num + den - 1) / den);
  I0 = q;
  td = double(num - q * den) / double(den);
  tf = float(td);
}

template <typename T, typename Ix = long>
__device__ __forceinline__ T coarse_at(const DLevel<T> &G, int dir, int I, int J, int K) {
  const bool out = I < 0 || I >= G.nx || J < 0 || J >= G.ny || K < 0 || K >= G.nz;
  if (out && dir) return T(0);
  I = min(max(I, 0), G.nx - 1);
  J = min(max(J, 0), G.ny - 1);
  K = min(max(K, 0), G.nz - 1);
  return G.sol[I + (Ix)G.nx * (J + (Ix)G.ny * K)];
}

template <typename T, int DIM, typename Ix = long>
__device__ __forceinline__ void prolong_cell(const DLevel<T> &G, const DLevel<T> &F, int r, int dir, Ix c) {
  int i, j, k;
  cell_ijk(c, F.nx, F.ny, i, j, k);
  int I0, J0, K0 = 0;
  float fx, fy, fz;
  double dx_, dy_, dz_ = 0.0;
  interp_axis(i, r, I0, fx, dx_);
  interp_axis(j, r, J0, fy, dy_);
  const T tx = sizeof(T) == 8 ? T(dx_) : T(fx), ty = sizeof(T) == 8 ? T(dy_) : T(fy);
  T v00 = (T(1) - tx) * coarse_at<T, Ix>(G, dir, I0, J0, K0) + tx * coarse_at<T, Ix>(G, dir, I0 + 1, J0, K0);
  T v10 = (T(1) - tx) * coarse_at<T, Ix>(G, dir, I0, J0 + 1, K0) + tx * coarse_at<T, Ix>(G, dir, I0 + 1, J0 + 1, K0);
  T v = (T(1) - ty) * v00 + ty * v10;
  if constexpr (DIM == 3) {
    interp_axis(k, r, K0, fz, dz_);
    const T tz = sizeof(T) == 8 ? T(dz_) : T(fz);
    v00 = (T(1) - tx) * coarse_at<T, Ix>(G, dir, I0, J0, K0) + tx * coarse_at<T, Ix>(G, dir, I0 + 1, J0, K0);
    v10 = (T(1) - tx) * coarse_at<T, Ix>(G, dir, I0, J0 + 1, K0) + tx * coarse_at<T, Ix>(G, dir, I0 + 1, J0 + 1, K0);
    const T v0 = (T(1) - ty) * v00 + ty * v10;
    const T v01 = (T(1) - tx) * coarse_at<T, Ix>(G, dir, I0, J0, K0 + 1) + tx * coarse_at<T, Ix>(G, dir, I0 + 1, J0, K0 + 1);
    const T v11 = (T(1) - tx) * coarse_at<T, Ix>(G, dir, I0, J0 + 1, K0 + 1) +
                  tx * coarse_at<T, Ix>(G, dir, I0 + 1, J0 + 1, K0 + 1);
    const T v1 = (T(1) - ty) * v01 + ty * v11;
    v = (T(1) - tz) * v0 + tz * v1;
  }
  F.sol[c] += v;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock) k_prolong_add(DLevel<T> G, DLevel<T> F, int r, int dir) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < F.n) prolong_cell<T, DIM>(G, F, r, dir, c);
}

// 3-D prolong-add for factor-2 coarsening: prolong_cell with r = 2.  Each
// thread owns the x-adjacent fine cells 2m, 2m+1 of one fine row; their x
// weights are 1/4, 3/4 on coarse cells m-1, m and 3/4, 1/4 on m, m+1, so a
// row pair reads three coarse values per coarse row and the fine pair moves
// as one vector.  Block: 64 pairs x 4 fine rows; grid z: fine planes.
constexpr int kP3X = 64, kP3Y = 4;
template <typename T>
__global__ void __launch_bounds__(kP3X *kP3Y) k_prolong3_r2(DLevel<T> G, DLevel<T> F, int dir) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const int m = blockIdx.x * kP3X + threadIdx.x;
  const int j = blockIdx.y * kP3Y + threadIdx.y, k = blockIdx.z;
  if (m >= G.nx || j >= F.ny) return;
  const int J0 = (j - 1) >> 1, K0 = (k - 1) >> 1;  // floor((2j - 1) / 4)
  const T ty = (j & 1) ? T(0.25) : T(0.75), tz = (k & 1) ? T(0.25) : T(0.75);
  T a[2][2][3];  // [J - J0][K - K0][I - m + 1], ghost rule of coarse_at
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int J = J0 + jj, K = K0 + kk;
      const bool rout = J < 0 || J >= G.ny || K < 0 || K >= G.nz;
      const T *row = G.sol + (long)G.nx * (min(max(J, 0), G.ny - 1) + (long)G.ny * min(max(K, 0), G.nz - 1));
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int I = m - 1 + d;
        const bool out = rout || I < 0 || I >= G.nx;
        a[jj][kk][d] = (out && dir) ? T(0) : row[min(max(I, 0), G.nx - 1)];
      }
    }
  auto interp = [&](T tx, int o) {
    const T v00 = (T(1) - tx) * a[0][0][o] + tx * a[0][0][o + 1];
    const T v10 = (T(1) - tx) * a[1][0][o] + tx * a[1][0][o + 1];
    const T v0 = (T(1) - ty) * v00 + ty * v10;
    const T v01 = (T(1) - tx) * a[0][1][o] + tx * a[0][1][o + 1];
    const T v11 = (T(1) - tx) * a[1][1][o] + tx * a[1][1][o + 1];
    const T v1 = (T(1) - ty) * v01 + ty * v11;
    return (T(1) - tz) * v0 + tz * v1;
  };
  // F.nx = 2 G.nx is even, so the pair is aligned to two elements
  T2 *p = reinterpret_cast<T2 *>(F.sol + 2 * m + (long)F.nx * (j + (long)F.ny * k));
  T2 f = *p;
  f.x += interp(T(0.75), 0);
  f.y += interp(T(0.25), 1);
  *p = f;
}

// --------------------------------------------- dense coarsest solve
// x = Minv b in FP64 on one workgroup; Neumann: subtract the mean (the
// pinned inverse maps a nonzero-mean b to x0 - mean(b)/s, see set-up).
template <typename T>
__device__ void coarse_solve_block(const DLevel<T> &G, const double *__restrict__ Minv, int neumann) {
  __shared__ double xs[kMaxCoarse];
  __shared__ double bs[kMaxCoarse];
  __shared__ double mean;
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
  const double tot = block_sum(local);
  if (threadIdx.x == 0) mean = neumann ? tot / n : 0.0;
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) G.sol[c] = T(xs[c] - mean);
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_coarse_solve(DLevel<T> G, const double *__restrict__ Minv,
                                                         int neumann) {
  coarse_solve_block(G, Minv, neumann);
}

// the one-workgroup kernels below run with exactly kBlockThreads threads; a
// constant stride keeps blockDim.x (a load from the dispatch packet) out of
// their loops
constexpr int kBlockThreads = 1024;
// k_block_vcycle's dense coarsest solve gives each thread at most two rows
static_assert(kMaxCoarse <= 2 * kBlockThreads, "coarsest rows exceed two per thread");

// LDS-resident levels hold < 2^31 cells: 32-bit indices throughout
template <typename T, int DIM, bool VAR>
__device__ void block_sweep(const DLevel<T> &L, int first, int zero_guess) {
  const int hx = (L.nx + 1) >> 1;
  const int half = hx * L.ny * L.nz;
  for (int pass = 0; pass < 2; ++pass) {
    const int color = pass == 0 ? first : 1 - first;
    for (int t = threadIdx.x; t < half; t += kBlockThreads) {
      int i, j, k;
      if constexpr (DIM == 2) {  // row j, colour cell t - j hx of it
        j = (int)((unsigned)t / (unsigned)hx);
        k = 0;
        i = 2 * (t - j * hx) + ((j + color) & 1);
        if (i >= L.nx) continue;
      } else {
        if (!color_cell(t, L.nx, L.ny, L.nz, color, i, j, k)) continue;
      }
      gs_cell<T, DIM, VAR, int>(L, i, j, k, pass == 0 ? zero_guess : 0);
    }
    __syncthreads();
  }
}

// ---------------------------------------- one-workgroup coarse V-cycle
// The latency-bound coarse levels (from the first level whose whole
// remaining hierarchy fits in the workgroup's LDS) are cycled by a single
// 1024-thread workgroup.  Every level's iterate, right-hand side and face
// coefficients are staged into LDS once; each phase (one colour of a sweep,
// restriction, dense coarsest solve, prolongation) is a strided loop over the
// level followed by a workgroup barrier, so phases cost LDS latency rather
// than global-memory round trips.  Only the top level's iterate is written
// back.
constexpr int kMaxBlockLevels = 12;

template <typename T>
struct BlockLevels {
  DLevel<T> lv[kMaxBlockLevels];  // global views (read/write-back)
  int ratio[kMaxBlockLevels];     // deg(l+1)/deg(l)
  int off[kMaxBlockLevels];       // element offset of level l in LDS
  int nlev;
  int arrays;                     // arrays per level in LDS (2 or 2 + ndim)
};

// LDS view of block level l: global metadata, LDS arrays.  Built from the
// extern __shared__ base on every use so the compiler keeps LDS (not flat)
// addressing.
template <typename T, bool VAR, int DIM>
__device__ __forceinline__ DLevel<T> lds_view(const BlockLevels<T> *__restrict__ B, int l, T *smem) {
  DLevel<T> v = B->lv[l];
  T *base = smem + B->off[l];
  const long n = v.n;
  v.sol = base;
  v.rhs = base + n;
  if constexpr (VAR) {
    v.cx = base + 2 * n;
    v.cy = base + 3 * n;
    if constexpr (DIM == 3) v.cz = base + 4 * n;
  }
  return v;
}

template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlockThreads)
    k_block_vcycle(const BlockLevels<T> *__restrict__ B, int nu1, int nu2, const double *__restrict__ Minv,
                   int neumann, int dir, int zero_guess0, int load_top) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int L = B->nlev;
  for (int l = 0; l < L; ++l) {
    const DLevel<T> g = B->lv[l];
    const DLevel<T> S = lds_view<T, VAR, DIM>(B, l, smem);
    for (long c = threadIdx.x; c < g.n; c += kBlockThreads) {
      S.sol[c] = (l == 0 && load_top && !zero_guess0) ? g.sol[c] : T(0);
      if (l == 0) S.rhs[c] = g.rhs[c];
      if constexpr (VAR) {
        const_cast<T *>(S.cx)[c] = g.cx[c];
        const_cast<T *>(S.cy)[c] = g.cy[c];
        if constexpr (DIM == 3) const_cast<T *>(S.cz)[c] = g.cz[c];
      }
    }
  }
  __syncthreads();
  for (int l = 0; l < L - 1; ++l) {
    const DLevel<T> F = lds_view<T, VAR, DIM>(B, l, smem);
    const DLevel<T> G = lds_view<T, VAR, DIM>(B, l + 1, smem);
    for (int s = 0; s < nu1; ++s) block_sweep<T, DIM, VAR>(F, 0, (l == 0 && s == 0) ? zero_guess0 : 0);
    const int r = B->ratio[l];
    for (int C = threadIdx.x; C < G.n; C += kBlockThreads) rr_cell<T, DIM, VAR, int>(F, G, r, C);
    __syncthreads();
  }
  {  // dense coarsest solve, FP64, rows in registers
    const DLevel<T> G = lds_view<T, VAR, DIM>(B, L - 1, smem);
    const int n = (int)G.n;
    double xr[2] = {0.0, 0.0};
    double local = 0.0;
    for (int q = 0; q < 2; ++q) {
      const int row = threadIdx.x + q * kBlockThreads;
      if (row < n) {
        const double *m = Minv + (long)row * n;
        double acc = 0.0;
        for (int c = 0; c < n; ++c) acc += m[c] * double(G.rhs[c]);
        xr[q] = acc;
        local += acc;
      }
    }
    const double tot = block_sum(local);
    __shared__ double mean;
    if (threadIdx.x == 0) mean = neumann ? tot / n : 0.0;
    __syncthreads();
    for (int q = 0; q < 2; ++q) {
      const int row = threadIdx.x + q * kBlockThreads;
      if (row < n) G.sol[row] = T(xr[q] - mean);
    }
    __syncthreads();
  }
  for (int l = L - 1; l > 0; --l) {
    const DLevel<T> G = lds_view<T, VAR, DIM>(B, l, smem);
    const DLevel<T> F = lds_view<T, VAR, DIM>(B, l - 1, smem);
    const int r = B->ratio[l - 1];
    for (int c = threadIdx.x; c < F.n; c += kBlockThreads) prolong_cell<T, DIM, int>(G, F, r, dir, c);
    __syncthreads();
    for (int s = 0; s < nu2; ++s) block_sweep<T, DIM, VAR>(F, 1, 0);
  }
  const DLevel<T> top = B->lv[0];
  const DLevel<T> S0 = lds_view<T, VAR, DIM>(B, 0, smem);
  for (long c = threadIdx.x; c < top.n; c += kBlockThreads) top.sol[c] = S0.sol[c];
}

// ---------------------------------------- fused 2-D fine-level kernels
// A workgroup owns a TX x TY tile of a fine level and stages the iterate,
// right-hand side and face coefficients of the tile plus a kFH-cell halo
// into LDS with one batch of independent loads.  Red-black sweeps then run
// on shrinking rings around the tile (a sweep consumes two rings: the first
// colour is updated one ring further out than the second), recomputing the
// ring cells that neighbouring tiles own from the same inputs, so tiles need
// no synchronisation with each other.  Iterates are written to a second
// buffer (neighbours always read the pre-kernel state).
//   k_pre2d : nu1 sweeps (red first), residual, child-average restriction to
//             the coarse level (coarse sol = 0), fine tile written.
//   k_post2d: coarse correction prolonged onto the whole staged region, nu2
//             sweeps (black first), fine tile written; MODE 1 also reduces the
//             squared residual of the tile (FP64 convergence test); MODE 2 is
//             the FP32 level 0 of the mixed solve: instead of writing the
//             correction e it forms p_new = p + e and the FP64 residual
//             b - A p_new of its tile, writing p_new, float(r) and the tile's
//             sum of r^2 (the outer defect-correction step, fused).
// Tiles are ordered so the workgroups that land on one XCD (round robin
// over 8) own a contiguous band of rows and share halos in that XCD's L2.
constexpr int kFH = 5;  // staged halo: 2*nu + 1 with nu <= 2
constexpr int kFusedThreads = 512;

__device__ __forceinline__ int xcd_tile(int ntiles) {
  const int b = blockIdx.x;
  if (ntiles % 8 != 0) return b;
  return (b % 8) * (ntiles / 8) + b / 8;
}

// blockIdx.x of a 1-D cell grid, remapped as xcd_tile does: each XCD gets
// a contiguous run of blocks, so a stencil's neighbours one row or plane
// away sit in the same L2.  Partial sums stay indexed by blockIdx.x.
__device__ __forceinline__ unsigned xcd_block() { return (unsigned)xcd_tile((int)gridDim.x); }

template <typename T, bool VAR, int TX, int TY>
struct Fused {
  static constexpr int W = TX + 2 * kFH, H = TY + 2 * kFH;
};

// stage x (unless zero), b, faces of the region into LDS
template <typename T, bool VAR, int TX, int TY>
__device__ __forceinline__ void fused_stage(const DLevel<T> &F, int x0, int y0, bool zero_x,
                                            T (*xs)[TX + 2 * kFH], T (*bs)[TX + 2 * kFH],
                                            T (*cxs)[TX + 2 * kFH], T (*cys)[TX + 2 * kFH]) {
  constexpr int W = TX + 2 * kFH, H = TY + 2 * kFH;
  for (int q = threadIdx.x; q < W * H; q += blockDim.x) {
    const int ly = q / W, lx = q % W;
    const int gx = x0 - kFH + lx, gy = y0 - kFH + ly;
    const bool in = gx >= 0 && gx < F.nx && gy >= 0 && gy < F.ny;
    const long g = gx + (long)F.nx * gy;
    xs[ly][lx] = (in && !zero_x) ? F.sol[g] : T(0);
    bs[ly][lx] = in ? F.rhs[g] : T(0);
    if constexpr (VAR) {
      cxs[ly][lx] = in ? F.cx[g] : T(0);
      cys[ly][lx] = in ? F.cy[g] : T(0);
    }
  }
}

// one colour on the tile grown by RING: only cells of that colour are
// enumerated (no idle lanes)
template <typename T, bool VAR, int TX, int TY, int RING>
__device__ __forceinline__ void fused_pass(const DLevel<T> &F, int x0, int y0, int color,
                                           T (*xs)[TX + 2 * kFH], T (*bs)[TX + 2 * kFH],
                                           T (*cxs)[TX + 2 * kFH], T (*cys)[TX + 2 * kFH]) {
  constexpr int w = TX + 2 * RING, h = TY + 2 * RING, hw = (w + 1) / 2;
  for (int q = threadIdx.x; q < h * hw; q += blockDim.x) {
    const int ry = q / hw;
    const int gy = y0 - RING + ry;
    // first column of this row with (gx + gy) % 2 == color
    const int rx = 2 * (q - ry * hw) + ((color - (x0 - RING) - gy) & 1);
    if (rx >= w) continue;
    const int gx = x0 - RING + rx;
    if (gx < 0 || gx >= F.nx || gy < 0 || gy >= F.ny) continue;
    const int lx = rx + kFH - RING, ly = ry + kFH - RING;
    T off, diag;
    T cL = T(0), cR = T(0), cD = T(0), cU = T(0);
    if constexpr (VAR) {
      cL = cxs[ly][lx - 1];
      cR = cxs[ly][lx];
      cD = cys[ly - 1][lx];
      cU = cys[ly][lx];
    }
    stencil2<T, VAR>(F, gx, gy, xs[ly][lx - 1], xs[ly][lx + 1], xs[ly - 1][lx], xs[ly + 1][lx], cL, cR, cD, cU, off,
                     diag);
    xs[ly][lx] = gs_div<T>(bs[ly][lx] - off, diag);
  }
}

// nu red-black sweeps (colour `first` first) on shrinking rings: the
// first colour of sweep s on ring 2 nu - 2 s, the second on 2 nu - 1 - 2 s
template <typename T, bool VAR, int TX, int TY>
__device__ __forceinline__ void fused_sweeps(const DLevel<T> &F, int x0, int y0, int nu, int first,
                                             T (*xs)[TX + 2 * kFH], T (*bs)[TX + 2 * kFH],
                                             T (*cxs)[TX + 2 * kFH], T (*cys)[TX + 2 * kFH]) {
  if (nu == 2) {
    fused_pass<T, VAR, TX, TY, 4>(F, x0, y0, first, xs, bs, cxs, cys);
    __syncthreads();
    fused_pass<T, VAR, TX, TY, 3>(F, x0, y0, 1 - first, xs, bs, cxs, cys);
    __syncthreads();
  }
  if (nu >= 1) {
    fused_pass<T, VAR, TX, TY, 2>(F, x0, y0, first, xs, bs, cxs, cys);
    __syncthreads();
    fused_pass<T, VAR, TX, TY, 1>(F, x0, y0, 1 - first, xs, bs, cxs, cys);
    __syncthreads();
  }
}

// residual of tile cell (lx, ly) from LDS
template <typename T, bool VAR, int TX>
__device__ __forceinline__ T fused_res(const DLevel<T> &F, int gx, int gy, int lx, int ly,
                                       T (*xs)[TX + 2 * kFH], T (*bs)[TX + 2 * kFH], T (*cxs)[TX + 2 * kFH],
                                       T (*cys)[TX + 2 * kFH]) {
  T off, diag;
  T cL = T(0), cR = T(0), cD = T(0), cU = T(0);
  if constexpr (VAR) {
    cL = cxs[ly][lx - 1];
    cR = cxs[ly][lx];
    cD = cys[ly - 1][lx];
    cU = cys[ly][lx];
  }
  stencil2<T, VAR>(F, gx, gy, xs[ly][lx - 1], xs[ly][lx + 1], xs[ly - 1][lx], xs[ly + 1][lx], cL, cR, cD, cU, off,
                   diag);
  return bs[ly][lx] - (off + diag * xs[ly][lx]);
}

template <typename T, bool VAR, int TX, int TY>
__global__ void __launch_bounds__(kFusedThreads)
    k_pre2d(DLevel<T> F, DLevel<T> G, T *__restrict__ out, int nu1, int zero_guess) {
  constexpr int W = TX + 2 * kFH, H = TY + 2 * kFH;
  __shared__ T xs[H][W];
  __shared__ T bs[H][W];
  __shared__ T cxs[VAR ? H : 1][W];
  __shared__ T cys[VAR ? H : 1][W];
  const int tiles_x = (F.nx + TX - 1) / TX, ntiles = tiles_x * ((F.ny + TY - 1) / TY);
  const int tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;
  const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
  fused_stage<T, VAR, TX, TY>(F, x0, y0, zero_guess != 0, xs, bs, cxs, cys);
  __syncthreads();
  fused_sweeps<T, VAR, TX, TY>(F, x0, y0, nu1, 0, xs, bs, cxs, cys);
  // residual + child average (ratio 2, tiles are even-aligned)
  for (int q = threadIdx.x; q < (TX / 2) * (TY / 2); q += blockDim.x) {
    const int I = x0 / 2 + q % (TX / 2), J = y0 / 2 + q / (TX / 2);
    if (I >= G.nx || J >= G.ny) continue;
    T sum = T(0);
    for (int jj = 0; jj < 2; ++jj)
      for (int ii = 0; ii < 2; ++ii) {
        const int gx = 2 * I + ii, gy = 2 * J + jj;
        sum += fused_res<T, VAR, TX>(F, gx, gy, gx - x0 + kFH, gy - y0 + kFH, xs, bs, cxs, cys);
      }
    const long C = I + (long)G.nx * J;
    G.rhs[C] = sum / T(4);
    G.sol[C] = T(0);
  }
  for (int q = threadIdx.x; q < TX * TY; q += blockDim.x) {
    const int gx = x0 + q % TX, gy = y0 + q / TX;
    if (gx < F.nx && gy < F.ny) out[gx + (long)F.nx * gy] = xs[gy - y0 + kFH][gx - x0 + kFH];
  }
}

// p.q and p.r from k_fcg_direction's partial sums (p.q at [0, nb), p.r
// after them), summed by the first wave in one fixed order, whatever the
// workgroup size, and returned to every thread: the FCG update and the fused
// update + pre-smoothing kernel form the same alpha.  corr_n > 0 (the cell
// count of a singular all-Neumann operator): the sums of p, r and q follow
// at [2 nb, 5 nb), and p.q, p.r become (p - mean p).q and (p - mean p).r.
// The constant component of p, which A maps to zero, then cannot enter the
// step sizes: left in, it grows once the residual is at rounding level and
// its product with the rounding-level sum of r drives the iteration away.
__device__ __forceinline__ void fcg_sums_of(double a, double b, double sp, double sr, double sq, double corr_n,
                                            double &pq, double &pr) {
  if (corr_n > 0.0) {
    a -= sp * sq / corr_n;
    b -= sp * sr / corr_n;
  }
  pq = a;
  pr = b;
}
// The partials reach wave 0 through LDS in chunks of kDirChunk, loaded by
// the whole workgroup at once (one memory round trip per chunk instead of
// one per partial and lane); lane l then adds partials l, l + kWave,
// l + 2 kWave, ... in that order, whatever the workgroup size.
constexpr int kDirChunk = 512;
__device__ __forceinline__ void fcg_dir_sums(const double *__restrict__ part, int nb, double corr_n, double &pq,
                                             double &pr) {
  __shared__ double ta, tb;
  __shared__ double buf[5][kDirChunk];
  const bool corr = corr_n > 0.0;
  double a = 0.0, b = 0.0, sp = 0.0, sr = 0.0, sq = 0.0;
  for (int c0 = 0; c0 < nb; c0 += kDirChunk) {
    for (int t = threadIdx.x; t < kDirChunk; t += blockDim.x) {
      const int k = c0 + t;
      const bool in = k < nb;
      buf[0][t] = in ? part[k] : 0.0;
      buf[1][t] = in ? part[nb + k] : 0.0;
      if (corr) {
        buf[2][t] = in ? part[2 * nb + k] : 0.0;
        buf[3][t] = in ? part[3 * nb + k] : 0.0;
        buf[4][t] = in ? part[4 * nb + k] : 0.0;
      }
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
      for (int t = threadIdx.x; t < kDirChunk && c0 + t < nb; t += kWave) {
        a += buf[0][t];
        b += buf[1][t];
        if (corr) {
          sp += buf[2][t];
          sr += buf[3][t];
          sq += buf[4][t];
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < kWave) {
    a = wave_sum(a);
    b = wave_sum(b);
    if (corr) {
      sp = wave_sum(sp);
      sr = wave_sum(sr);
      sq = wave_sum(sq);
    }
    if (threadIdx.x == 0) fcg_sums_of(a, b, sp, sr, sq, corr_n, ta, tb);
  }
  __syncthreads();
  pq = ta;
  pr = tb;
}

// The FCG update of the mixed solve fused into the next V-cycle's level-0
// pre-smoothing (k_pre2d on the FP32 level 0 from a zero guess).  alpha =
// p.r / p.q from the direction kernel's partial sums; the staged right-hand
// side is float(r - alpha q), formed for the halo cells too, from the
// residual before the update (r_in, read-only here); the tile's own cells
// get x += alpha p, r_out = r - alpha q, the FP32 level-0 right-hand side
// float(r_out) and a per-tile sum of r_out^2.  Workgroup 0 stores p.q and
// p.r for the next beta.  Every value is the one k_fcg_update followed by
// k_pre2d would produce.
struct FcgUpdate {
  double *x;
  const double *r_in;
  double *r_out;
  const double *p, *q;
  double *scal;
  const double *dir_part;
  int dir_nb;
  double corr_n;    // fcg_dir_sums' null-space correction (0: none)
  double *r2_part;  // indexed by workgroup
  // measurement (nullable): workgroup b stores its start and end on the
  // device wall clock in clk[2 b], clk[2 b + 1] (the kernel's own span is
  // the earliest start to the latest end)
  unsigned long long *clk;
};

template <bool VAR, int TX, int TY>
__global__ void __launch_bounds__(kFusedThreads)
    k_pre2d_fcg(DLevel<float> F, DLevel<float> G, float *__restrict__ out, int nu1, FcgUpdate u) {
  constexpr int W = TX + 2 * kFH, H = TY + 2 * kFH;
  __shared__ float xs[H][W];
  __shared__ float bs[H][W];
  __shared__ float cxs[VAR ? H : 1][W];
  __shared__ float cys[VAR ? H : 1][W];
  const int tiles_x = (F.nx + TX - 1) / TX, ntiles = tiles_x * ((F.ny + TY - 1) / TY);
  const int tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;  // never taken: the grid has exactly ntiles blocks
  const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
  unsigned long long t_start = 0;
  if (u.clk && threadIdx.x == 0) t_start = (unsigned long long)wall_clock64();
  // r and q of the staged region are loaded before alpha's reduction (the
  // two latencies overlap), then combined once alpha is known; the largest
  // tile loads them afterwards (its registers would cost occupancy)
  // (the small tiles also load x and p of their own cells there, and
  // update them in the staging pass, so the tile is not read twice)
  constexpr int NS0 = (W * H + kFusedThreads - 1) / kFusedThreads;
  constexpr bool early = NS0 <= 6;
  constexpr int NS = early ? NS0 : 1;
  double rv[NS], qv[NS], xv[NS], pv[NS];
  auto in_tile = [&](int lx, int ly, int gx, int gy) {
    return lx >= kFH && lx < kFH + TX && ly >= kFH && ly < kFH + TY && gx < F.nx && gy < F.ny;
  };
#pragma unroll
  for (int t = 0; t < NS0; ++t) {
    const int q = threadIdx.x + t * kFusedThreads;
    if (q < W * H) {
      const int ly = q / W, lx = q % W;
      const int gx = x0 - kFH + lx, gy = y0 - kFH + ly;
      const bool in = gx >= 0 && gx < F.nx && gy >= 0 && gy < F.ny;
      const long g = gx + (long)F.nx * gy;
      if constexpr (early) {
        rv[t] = in ? u.r_in[g] : 0.0;
        qv[t] = in ? u.q[g] : 0.0;
        const bool own = in_tile(lx, ly, gx, gy);
        xv[t] = own ? u.x[g] : 0.0;
        pv[t] = own ? u.p[g] : 0.0;
      }
      xs[ly][lx] = 0.0f;
      if constexpr (VAR) {
        cxs[ly][lx] = in ? F.cx[g] : 0.0f;
        cys[ly][lx] = in ? F.cy[g] : 0.0f;
      }
    }
  }
  double pq, pr;
  fcg_dir_sums(u.dir_part, u.dir_nb, u.corr_n, pq, pr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    u.scal[S_PQ] = pq;
    u.scal[S_PR] = pr;
  }
  const double alpha = pq != 0.0 ? pr / pq : 0.0;  // 0 only after exact convergence
  double r2 = 0.0;
#pragma unroll
  for (int t = 0; t < NS0; ++t) {
    const int q = threadIdx.x + t * kFusedThreads;
    if (q < W * H) {
      const int ly = q / W, lx = q % W;
      const int gx = x0 - kFH + lx, gy = y0 - kFH + ly;
      const bool in = gx >= 0 && gx < F.nx && gy >= 0 && gy < F.ny;
      const long g = gx + (long)F.nx * gy;
      if constexpr (early) {
        const double rc = rv[t] - alpha * qv[t];
        bs[ly][lx] = in ? float(rc) : 0.0f;
        if (in_tile(lx, ly, gx, gy)) {
          u.x[g] = xv[t] + alpha * pv[t];
          u.r_out[g] = rc;
          F.rhs[g] = float(rc);
          r2 += rc * rc;
        }
      } else {
        bs[ly][lx] = in ? float(u.r_in[g] - alpha * u.q[g]) : 0.0f;
      }
    }
  }
  if constexpr (!early) {
    for (int q = threadIdx.x; q < TX * TY; q += blockDim.x) {
      const int gx = x0 + q % TX, gy = y0 + q / TX;
      if (gx < F.nx && gy < F.ny) {
        const long g = gx + (long)F.nx * gy;
        u.x[g] += alpha * u.p[g];
        const double rc = u.r_in[g] - alpha * u.q[g];
        u.r_out[g] = rc;
        F.rhs[g] = float(rc);
        r2 += rc * rc;
      }
    }
  }
  r2 = block_sum(r2);  // its barriers also order the staging above
  if (threadIdx.x == 0) u.r2_part[blockIdx.x] = r2;
  fused_sweeps<float, VAR, TX, TY>(F, x0, y0, nu1, 0, xs, bs, cxs, cys);
  // residual + child average (ratio 2, tiles are even-aligned)
  for (int q = threadIdx.x; q < (TX / 2) * (TY / 2); q += blockDim.x) {
    const int I = x0 / 2 + q % (TX / 2), J = y0 / 2 + q / (TX / 2);
    if (I >= G.nx || J >= G.ny) continue;
    float sum = 0.0f;
    for (int jj = 0; jj < 2; ++jj)
      for (int ii = 0; ii < 2; ++ii) {
        const int gx = 2 * I + ii, gy = 2 * J + jj;
        sum += fused_res<float, VAR, TX>(F, gx, gy, gx - x0 + kFH, gy - y0 + kFH, xs, bs, cxs, cys);
      }
    const long C = I + (long)G.nx * J;
    G.rhs[C] = sum / 4.0f;
    G.sol[C] = 0.0f;
  }
  for (int q = threadIdx.x; q < TX * TY; q += blockDim.x) {
    const int gx = x0 + q % TX, gy = y0 + q / TX;
    if (gx < F.nx && gy < F.ny) out[gx + (long)F.nx * gy] = xs[gy - y0 + kFH][gx - x0 + kFH];
  }
  if (u.clk) {
    __syncthreads();
    if (threadIdx.x == 0) {
      u.clk[2 * blockIdx.x] = t_start;
      u.clk[2 * blockIdx.x + 1] = (unsigned long long)wall_clock64();
    }
  }
}

struct OuterArgs {  // level-0 extras of the post-smoothing kernel
  DLevel<double> P;         // MODE 2: rhs = b, faces / constants of the FP64 operator
  const double *p_old;      // MODE 2: current FP64 iterate
  double *p_new;            // MODE 2: updated FP64 iterate
  float *r32;               // MODE 2: float(b - A p_new) -> next FP32 right-hand side
  const double *q;          // MODE 3: previous search direction image A p
};

template <typename T, bool VAR, int TX, int TY, int MODE>
__global__ void __launch_bounds__(kFusedThreads)
    k_post2d(DLevel<T> G, DLevel<T> F, T *__restrict__ out, int nu2, int dir, double *__restrict__ partial,
             OuterArgs oa) {
  constexpr int W = TX + 2 * kFH, H = TY + 2 * kFH;
  constexpr int CW = TX / 2 + 8, CH = TY / 2 + 8;
  __shared__ T xs[H][W];
  __shared__ T bs[H][W];
  __shared__ T cxs[VAR ? H : 1][W];
  __shared__ T cys[VAR ? H : 1][W];
  __shared__ T cs[CH][CW];
  const int tiles_x = (F.nx + TX - 1) / TX, ntiles = tiles_x * ((F.ny + TY - 1) / TY);
  const int tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;  // never taken: the grid has exactly ntiles blocks
  const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
  const int cx0 = x0 / 2 - 4, cy0 = y0 / 2 - 4;
  fused_stage<T, VAR, TX, TY>(F, x0, y0, false, xs, bs, cxs, cys);
  for (int q = threadIdx.x; q < CW * CH; q += blockDim.x) {
    const int ly = q / CW, lx = q % CW;
    int I = cx0 + lx, J = cy0 + ly;
    const bool out_c = I < 0 || I >= G.nx || J < 0 || J >= G.ny;
    T v = T(0);
    if (!(out_c && dir)) {
      I = min(max(I, 0), G.nx - 1);
      J = min(max(J, 0), G.ny - 1);
      v = G.sol[I + (long)G.nx * J];
    }
    cs[ly][lx] = v;
  }
  __syncthreads();
  // prolongation onto the whole staged region (cell-centred linear, ratio 2)
  for (int q = threadIdx.x; q < W * H; q += blockDim.x) {
    const int ly = q / W, lx = q % W;
    const int gx = x0 - kFH + lx, gy = y0 - kFH + ly;
    if (gx < 0 || gx >= F.nx || gy < 0 || gy >= F.ny) continue;
    int I0, J0;
    float tfx, tfy;
    double tdx, tdy;
    interp_axis(gx, 2, I0, tfx, tdx);
    interp_axis(gy, 2, J0, tfy, tdy);
    const T tx = sizeof(T) == 8 ? T(tdx) : T(tfx), ty = sizeof(T) == 8 ? T(tdy) : T(tfy);
    const int a = I0 - cx0, b = J0 - cy0;
    const T v00 = (T(1) - tx) * cs[b][a] + tx * cs[b][a + 1];
    const T v10 = (T(1) - tx) * cs[b + 1][a] + tx * cs[b + 1][a + 1];
    xs[ly][lx] += (T(1) - ty) * v00 + ty * v10;
  }
  __syncthreads();
  fused_sweeps<T, VAR, TX, TY>(F, x0, y0, nu2, 1, xs, bs, cxs, cys);
  double local = 0.0;
  for (int q = threadIdx.x; q < TX * TY; q += blockDim.x) {
    const int gx = x0 + q % TX, gy = y0 + q / TX;
    if (gx >= F.nx || gy >= F.ny) continue;
    const int lx = gx - x0 + kFH, ly = gy - y0 + kFH;
    const long g = gx + (long)F.nx * gy;
    if constexpr (MODE == 0) {
      out[g] = xs[ly][lx];
    } else if constexpr (MODE == 3) {
      out[g] = xs[ly][lx];
      local += double(xs[ly][lx]) * oa.q[g];
    } else if constexpr (MODE == 1) {
      out[g] = xs[ly][lx];
      const double r = double(fused_res<T, VAR, TX>(F, gx, gy, lx, ly, xs, bs, cxs, cys));
      local += r * r;
    } else {
      // p_new = p_old + e on the cell and its neighbours, FP64 residual
      const DLevel<double> &P = oa.P;
      const long nx = F.nx;
      auto pv = [&](long gg, int llx, int lly) { return oa.p_old[gg] + double(xs[lly][llx]); };
      const double pc = pv(g, lx, ly);
      const double pl = gx > 0 ? pv(g - 1, lx - 1, ly) : 0.0;
      const double pr = gx < F.nx - 1 ? pv(g + 1, lx + 1, ly) : 0.0;
      const double pd = gy > 0 ? pv(g - nx, lx, ly - 1) : 0.0;
      const double pu = gy < F.ny - 1 ? pv(g + nx, lx, ly + 1) : 0.0;
      double cL = 0.0, cR = 0.0, cD = 0.0, cU = 0.0;
      if constexpr (VAR) {
        cL = gx > 0 ? P.cx[g - 1] : 0.0;
        cR = P.cx[g];
        cD = gy > 0 ? P.cy[g - nx] : 0.0;
        cU = P.cy[g];
      }
      double off, diag;
      stencil2<double, VAR>(P, gx, gy, pl, pr, pd, pu, cL, cR, cD, cU, off, diag);
      const double r = P.rhs[g] - (off + diag * pc);
      oa.p_new[g] = pc;
      oa.r32[g] = float(r);
      local += r * r;
    }
  }
  if constexpr (MODE != 0) {
    const double tot = block_sum(local);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
  }
}

// ======================================================= 3-D z-streaming
// One red-black Gauss-Seidel sweep of a 3-D level per launch.  A workgroup
// owns a TX x TY column tile and a chunk of z planes and marches up the
// chunk with a ring of five planes in LDS (the tile plus a 2-cell ring) of
// x and the coefficient source.  Step `pl` stages plane pl; then, in one
// phase, it updates the first colour on plane pl-1 over the tile and a
// 1-cell ring (ring cells are computed redundantly by the neighbouring
// tiles, from the same inputs) and the second colour on plane pl-3 over the
// tile -- the two read disjoint planes of the other's colour, so they need
// no barrier between them -- and stores plane pl-3, now final, to `out`.
// `in` is never written, so tiles and chunks never read each other's new
// values.  Planes pl+1 and pl+2 are in flight in registers, and the right-
// hand sides of the next step's cells are loaded one step ahead straight
// into registers (b is read once per updated cell, so it needs no LDS).
// A chunk also updates the first colour on the plane below and above it
// (redundantly), which its second colour needs.
// COEF 0: constant operator (ax, ay, az; Dirichlet wall terms as stencil());
// COEF 1: face arrays cx, cy, cz; COEF 2: density rho with the face formed
// on the fly, the level-0 operator of init_ppe_proj_coefficients
// (projection_method.cpp:404-465); cells outside the domain hold rho = +inf,
// so wall faces vanish.
constexpr int kZ3Threads = 256;
constexpr int kZ3Ring = 5;

template <int COEF>
struct Z3 {
  static constexpr int NA = COEF == 1 ? 4 : (COEF == 2 ? 2 : 1);  // x + coefficient arrays
};

__device__ __forceinline__ int z3_slot(int plane) { return (plane + 2 * kZ3Ring) % kZ3Ring; }

// face between cells of density a and b: 2 / (a + b) / h^2, equal to
// init_ppe_proj_coefficients' 1 / (0.5 (a + b)) / h / h
// (projection_method.cpp:404-465) up to rounding; inv_h2 = 1 / h^2.  A
// neighbour outside the domain carries rho = +inf and gives 0.
template <typename T>
__device__ __forceinline__ T rho_face(T a, T b, T inv_h2) {
  return inv_h2 * (T(2) * smooth_rcp<T>(a + b));
}

// off-diagonal sum and diagonal of cell (gx, gy, gz) from LDS neighbours;
// same accumulation order as stencil()
template <typename T, int COEF>
__device__ __forceinline__ void z3_stencil(const DLevel<T> &F, int gx, int gy, int gz, T xl, T xr, T xd, T xu,
                                           T xb, T xf, T cL, T cR, T cD, T cU, T cB, T cF, T &off, T &diag) {
  T o = T(0);
  if constexpr (COEF != 0) {
    if (gx > 0) o += cL * xl;
    if (gx < F.nx - 1) o += cR * xr;
    if (gy > 0) o += cD * xd;
    if (gy < F.ny - 1) o += cU * xu;
    T d = -(cR + cL + cU + cD);
    if (gz > 0) o += cB * xb;
    if (gz < F.nz - 1) o += cF * xf;
    d -= (cF + cB);
    diag = d;
  } else {
    T cc = T(0);
    if (gx > 0) { o += F.ax * xl; cc += F.ax; } else if (F.rbwall) cc += F.wx;
    if (gx < F.nx - 1) { o += F.ax * xr; cc += F.ax; } else if (F.rbwall) cc += F.wx;
    if (gy > 0) { o += F.ay * xd; cc += F.ay; } else if (F.rbwall) cc += F.wy;
    if (gy < F.ny - 1) { o += F.ay * xu; cc += F.ay; } else if (F.rbwall) cc += F.wy;
    if (gz > 0) { o += F.az * xb; cc += F.az; } else if (F.rbwall) cc += F.wz;
    if (gz < F.nz - 1) { o += F.az * xf; cc += F.az; } else if (F.rbwall) cc += F.wz;
    diag = (F.dirichlet && !F.rbwall) ? F.dconst : -cc;
  }
  off = o;
}

template <typename T, int COEF, int TX, int TY>
__global__ void __launch_bounds__(kZ3Threads)
    k_sweep3d(DLevel<T> F, const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ rho, T hx, T hy,
              T hz, int first, int zero_guess, int zchunk) {
  constexpr int W = TX + 4, H = TY + 4, P = W * H;
  constexpr int NA = Z3<COEF>::NA;
  constexpr int PER = (P + kZ3Threads - 1) / kZ3Threads;
  // first-colour cells: tile + 1 ring, half of each row; second colour: tile
  constexpr int RW = TX + 2, RH = TY + 2, HW0 = (RW + 1) / 2, N0 = RH * HW0;
  constexpr int HW1 = (TX + 1) / 2, N1 = TY * HW1;
  constexpr int M0 = (N0 + kZ3Threads - 1) / kZ3Threads, M1 = (N1 + kZ3Threads - 1) / kZ3Threads;
  __shared__ T sm[NA][kZ3Ring][P];
  const int tiles_x = (F.nx + TX - 1) / TX;
  const int ntiles = tiles_x * ((F.ny + TY - 1) / TY);
  const int nblocks = ntiles * ((F.nz + zchunk - 1) / zchunk);
  const int b = xcd_tile(nblocks);
  const int tile = b % ntiles, chunk = b / ntiles;
  const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
  const int z0 = chunk * zchunk, z1 = min(z0 + zchunk, F.nz);
  const int c0 = first, c1 = 1 - first;
  const long sxy = (long)F.nx * F.ny;

  auto fetch = [&](T (&reg)[NA][PER], int pl) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int q = threadIdx.x + r * kZ3Threads;
      const int ly = q / W, lx = q - ly * W;
      const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
      const bool inside = q < P && pl >= 0 && pl < F.nz && gx >= 0 && gx < F.nx && gy >= 0 && gy < F.ny;
      const long g = inside ? gx + (long)F.nx * gy + sxy * pl : 0;
      reg[0][r] = (inside && !zero_guess) ? in[g] : T(0);
      if constexpr (COEF == 1) {
        reg[1][r] = inside ? F.cx[g] : T(0);
        reg[2][r] = inside ? F.cy[g] : T(0);
        reg[3][r] = inside ? F.cz[g] : T(0);
      }
      if constexpr (COEF == 2) reg[1][r] = inside ? rho[g] : T(INFINITY);
    }
  };
  auto commit = [&](const T (&reg)[NA][PER], int slot) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int q = threadIdx.x + r * kZ3Threads;
      if (q < P) {
#pragma unroll
        for (int a = 0; a < NA; ++a) sm[a][slot][q] = reg[a][r];
      }
    }
  };
  // m-th first-colour cell of this thread on plane k (false if none)
  auto cell0 = [&](int m, int k, int &gx, int &gy, int &q) {
    const int t = threadIdx.x + m * kZ3Threads;
    const int ry = t / HW0;
    gy = y0 - 1 + ry;
    const int rx = 2 * (t - ry * HW0) + ((c0 - (x0 - 1) - gy - k) & 1);
    gx = x0 - 1 + rx;
    q = (ry + 1) * W + rx + 1;
    return t < N0 && rx < RW && gx >= 0 && gx < F.nx && gy >= 0 && gy < F.ny;
  };
  // m-th second-colour cell of this thread on plane k (false if none)
  auto cell1 = [&](int m, int k, int &gx, int &gy, int &q) {
    const int t = threadIdx.x + m * kZ3Threads;
    const int ry = t / HW1;
    gy = y0 + ry;
    const int rx = 2 * (t - ry * HW1) + ((c1 - x0 - gy - k) & 1);
    gx = x0 + rx;
    q = (ry + 2) * W + rx + 2;
    return t < N1 && rx < TX && gx < F.nx && gy < F.ny;
  };
  auto first_ok = [&](int k) { return k >= z0 - 1 && k <= z1 && k >= 0 && k < F.nz; };
  auto second_ok = [&](int k) { return k >= z0 && k < z1; };
  // right-hand sides of the cells a step updates (first colour on k0,
  // second on k1), loaded ahead into registers
  auto load_rhs = [&](T (&r0)[M0], T (&r1)[M1], int k0, int k1) {
#pragma unroll
    for (int m = 0; m < M0; ++m) {
      int gx, gy, q;
      r0[m] = (first_ok(k0) && cell0(m, k0, gx, gy, q)) ? F.rhs[gx + (long)F.nx * gy + sxy * k0] : T(0);
    }
#pragma unroll
    for (int m = 0; m < M1; ++m) {
      int gx, gy, q;
      r1[m] = (second_ok(k1) && cell1(m, k1, gx, gy, q)) ? F.rhs[gx + (long)F.nx * gy + sxy * k1] : T(0);
    }
  };
  auto update = [&](int q, int sc, int sb, int sf, int gx, int gy, int gz, T bq) {
    T cL = T(0), cR = T(0), cD = T(0), cU = T(0), cB = T(0), cF = T(0);
    if constexpr (COEF == 1) {
      cL = sm[1][sc][q - 1];
      cR = sm[1][sc][q];
      cD = sm[2][sc][q - W];
      cU = sm[2][sc][q];
      cB = sm[3][sb][q];
      cF = sm[3][sc][q];
    }
    if constexpr (COEF == 2) {
      const T rc = sm[1][sc][q];
      cL = rho_face<T>(sm[1][sc][q - 1], rc, F.ax);
      cR = rho_face<T>(rc, sm[1][sc][q + 1], F.ax);
      cD = rho_face<T>(sm[1][sc][q - W], rc, F.ay);
      cU = rho_face<T>(rc, sm[1][sc][q + W], F.ay);
      cB = rho_face<T>(sm[1][sb][q], rc, F.az);
      cF = rho_face<T>(rc, sm[1][sf][q], F.az);
    }
    T off, diag;
    z3_stencil<T, COEF>(F, gx, gy, gz, sm[0][sc][q - 1], sm[0][sc][q + 1], sm[0][sc][q - W], sm[0][sc][q + W],
                        sm[0][sb][q], sm[0][sf][q], cL, cR, cD, cU, cB, cF, off, diag);
    sm[0][sc][q] = gs_div<T>(bq - off, diag);
  };

  const int zs = z0 - 2;  // first plane staged
  const int zl = z1 + 1;  // last plane staged
  const int ze = z1 + 2;  // last step (second colour on z1 - 1)
  auto step = [&](int pl, T (&reg)[NA][PER], const T (&r0)[M0], const T (&r1)[M1], T (&n0)[M0], T (&n1)[M1]) {
    if (pl <= zl) commit(reg, z3_slot(pl));
    __syncthreads();
    if (pl + 2 <= zl) fetch(reg, pl + 2);
    load_rhs(n0, n1, pl, pl - 2);  // the next step's cells
    const int k = pl - 1, k3 = pl - 3;
    if (first_ok(k)) {
#pragma unroll
      for (int m = 0; m < M0; ++m) {
        int gx, gy, q;
        if (cell0(m, k, gx, gy, q)) update(q, z3_slot(k), z3_slot(k - 1), z3_slot(k + 1), gx, gy, k, r0[m]);
      }
    }
    if (second_ok(k3)) {
#pragma unroll
      for (int m = 0; m < M1; ++m) {
        int gx, gy, q;
        if (cell1(m, k3, gx, gy, q)) update(q, z3_slot(k3), z3_slot(k3 - 1), z3_slot(k3 + 1), gx, gy, k3, r1[m]);
      }
    }
    __syncthreads();
    if (second_ok(k3)) {
      const int sc = z3_slot(k3);
      for (int t = threadIdx.x; t < TX * TY; t += kZ3Threads) {
        const int ry = t / TX, rx = t - ry * TX;
        const int gx = x0 + rx, gy = y0 + ry;
        if (gx < F.nx && gy < F.ny) out[gx + (long)F.nx * gy + sxy * k3] = sm[0][sc][(ry + 2) * W + rx + 2];
      }
    }
  };
  T reg0[NA][PER], reg1[NA][PER];
  T ra0[M0], ra1[M1], rb0[M0], rb1[M1];
  fetch(reg0, zs);
  fetch(reg1, zs + 1);
  load_rhs(ra0, ra1, zs - 1, zs - 3);
  for (int pl = zs; pl <= ze; pl += 2) {
    step(pl, reg0, ra0, ra1, rb0, rb1);
    if (pl + 1 > ze) break;
    step(pl + 1, reg1, rb0, rb1, ra0, ra1);
  }
}

// coarse.rhs = child average of (b - A x), coarse.sol = 0 for factor-2
// coarsening of a 3-D z-streaming level, with the 2 x 2 x 2 children of a
// coarse cell spread over four threads: thread (m, jj, kk) of a 64 x 2 x 2
// block takes the x-adjacent fine cells 2m, 2m+1 of fine row (2J + jj,
// 2K + kk), and the block sums the four pair sums in LDS.  The residual of
// a cell is k_rr3d's; only the order of the eight-term sum differs.
constexpr int kRR3X = 64;
template <typename T, int COEF>
__global__ void __launch_bounds__(kRR3X * 4) k_rr3d_pairs(DLevel<T> F, DLevel<T> G, const T *__restrict__ rho) {
  __shared__ T part[4][kRR3X];
  const int m = blockIdx.x * kRR3X + threadIdx.x;
  const int J = blockIdx.y, K = blockIdx.z;
  const int j = 2 * J + threadIdx.y, k = 2 * K + threadIdx.z;
  const long sy = F.nx, sz = (long)F.nx * F.ny;
  T s = T(0);
  if (m < G.nx) {
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
      const int i = 2 * m + ii;
      const long c = i + sy * j + sz * k;
      T off, diag;
      if constexpr (COEF == 2) {
        const T inf = T(INFINITY), rc = rho[c];
        const T cL = rho_face<T>(i > 0 ? rho[c - 1] : inf, rc, F.ax);
        const T cR = rho_face<T>(rc, i < F.nx - 1 ? rho[c + 1] : inf, F.ax);
        const T cD = rho_face<T>(j > 0 ? rho[c - sy] : inf, rc, F.ay);
        const T cU = rho_face<T>(rc, j < F.ny - 1 ? rho[c + sy] : inf, F.ay);
        const T cB = rho_face<T>(k > 0 ? rho[c - sz] : inf, rc, F.az);
        const T cF = rho_face<T>(rc, k < F.nz - 1 ? rho[c + sz] : inf, F.az);
        const T *x = F.sol;
        z3_stencil<T, 2>(F, i, j, k, i > 0 ? x[c - 1] : T(0), i < F.nx - 1 ? x[c + 1] : T(0),
                         j > 0 ? x[c - sy] : T(0), j < F.ny - 1 ? x[c + sy] : T(0), k > 0 ? x[c - sz] : T(0),
                         k < F.nz - 1 ? x[c + sz] : T(0), cL, cR, cD, cU, cB, cF, off, diag);
      } else {
        stencil<T, 3, COEF == 1>(F, F.sol, i, j, k, c, off, diag);
      }
      s += F.rhs[c] - (off + diag * F.sol[c]);
    }
  }
  const int w = threadIdx.y + 2 * threadIdx.z;
  part[w][threadIdx.x] = s;
  __syncthreads();
  if (w == 0 && m < G.nx) {
    const long C = m + (long)G.nx * (J + (long)G.ny * K);
    G.rhs[C] = (((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) + part[3][threadIdx.x]) /
               T(8);
    G.sol[C] = T(0);
  }
}

// --------------------------------------------------- residual + norm
// res = b - A x (stored when store != 0); partial[block] = sum res^2.
template <typename T, int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_residual(DLevel<T> L, int store, double *__restrict__ partial,
                                                     const double *__restrict__ flag) {
  if (flag && *flag != 0.0) return;
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
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
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
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

// out[b] = sum of in[b*group .. (b+1)*group) (first stage of a large
// reduction; fixed order, so results are deterministic)
__global__ void __launch_bounds__(kBlock) k_compress(const double *__restrict__ in, int n, int group,
                                                     double *__restrict__ out) {
  const int b0 = blockIdx.x * group, b1 = min(n, b0 + group);
  double v = 0.0;
  for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) v += in[i];
  v = block_sum(v);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// scal[s0] = sum partial[0..nb), scal[s1] = sum partial[nb..2nb)
__global__ void k_finalize2(const double *__restrict__ partial, int nb, double *__restrict__ scal, int s0, int s1) {
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) {
    a += partial[q];
    b += partial[nb + q];
  }
  a = block_sum(a);
  b = block_sum(b);
  if (threadIdx.x == 0) {
    scal[s0] = a;
    scal[s1] = b;
  }
}

// scal[S_PQ], scal[S_PR] from the direction kernel's nb partial sums (two
// sets, or five with the null-space correction of fcg_dir_sums)
__global__ void __launch_bounds__(1024) k_fcg_dir_final(const double *__restrict__ partial, int nb,
                                                        double *__restrict__ scal, double corr_n) {
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int sets = corr_n > 0.0 ? 5 : 2;
  for (int q = threadIdx.x; q < nb; q += blockDim.x)
    for (int s = 0; s < sets; ++s) v[s] += partial[s * nb + q];
  block_sum_n<5>(v);
  if (threadIdx.x == 0) fcg_sums_of(v[0], v[1], v[2], v[3], v[4], corr_n, scal[S_PQ], scal[S_PR]);
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
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
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

// ------------------------------------------- flexible CG around the V-cycle
// FCG(1) (Notay): the V-cycle (FP64, or FP32 in the mixed solve) is the
// preconditioner z = B r of a conjugate-gradient iteration whose vectors
// x, r, p, q = A p and inner products stay FP64.  The preconditioner may be
// non-symmetric and inexact, hence the flexible direction update
// beta = -(z . q_old) / (p_old . q_old).  All scalars stay on the device.

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dot_zq(const T *__restrict__ z, const double *__restrict__ q, long n,
                                                   double *__restrict__ partial) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = c < n ? double(z[c]) * q[c] : 0.0;
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// Workgroup size of the two FCG vector kernels: every workgroup reduces
// the producer's partial sums itself, so fewer, larger workgroups read
// fewer of them (512 x 1024 threads at 512 x 1024 cells).
constexpr int kFcgThreads = 1024;

// p = z + beta p_old (beta = 0 on the first iteration), q = A p, partial sums
// of p.q and p.r.  beta = -(z.q_old)/(p_old.q_old): z.q_old from scal[S_ZQ],
// or, with zq_n > 0, reduced here from the post-smoothing kernel's zq_n
// partials (every workgroup sums them in the same order, so all get the
// same value), which saves the separate reduction launch.
template <typename T, int DIM, bool VAR, bool SUMS>
__global__ void __launch_bounds__(kFcgThreads)
    k_fcg_direction(DLevel<double> L, const T *__restrict__ z, const double *__restrict__ p_old,
                    double *__restrict__ p, double *__restrict__ q, const double *__restrict__ r,
                    const double *__restrict__ scal, int first, double *__restrict__ partial,
                    const double *__restrict__ zq_part, int zq_n) {
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
  double beta = 0.0;
  if (!first) {
    double zq;
    if (zq_n > 0) {
      double s = 0.0;
      for (int b = threadIdx.x; b < zq_n; b += blockDim.x) s += zq_part[b];
      zq = block_allsum(s);
    } else {
      zq = scal[S_ZQ];
    }
    const double pq = scal[S_PQ];
    beta = pq != 0.0 ? -zq / pq : 0.0;  // 0 only after exact convergence
  }
  double vpq = 0.0, vpr = 0.0, s3[3] = {0.0, 0.0, 0.0};
  if (c < L.n) {
    int i, j, k;
    cell_ijk(c, L.nx, L.ny, i, j, k);
    const long sy = L.nx, sz = (long)L.nx * L.ny;
    auto P = [&](long g) { return double(z[g]) + (first ? 0.0 : beta * p_old[g]); };
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
    const double qc = off + diag * pc;
    const double rc = r[c];
    p[c] = pc;
    q[c] = qc;
    vpq = pc * qc;
    vpr = pc * rc;
    if constexpr (SUMS) {
      s3[0] = pc;
      s3[1] = rc;
      s3[2] = qc;
    }
  }
  vpq = block_sum(vpq);
  vpr = block_sum(vpr);
  if constexpr (SUMS) block_sum_n<3>(s3);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = vpq;
    partial[gridDim.x + blockIdx.x] = vpr;
    if constexpr (SUMS) {
      partial[2 * gridDim.x + blockIdx.x] = s3[0];
      partial[3 * gridDim.x + blockIdx.x] = s3[1];
      partial[4 * gridDim.x + blockIdx.x] = s3[2];
    }
  }
}

// 3-D k_fcg_direction marching in z: thread t of a kZmX x kZmY block owns
// column (i, j) = (bx kZmX + t % kZmX, by kZmY + t / kZmX) of a chunk of
// planes.  p = z + beta p_old of the plane below and above comes from a
// register window (formed once per cell, as k_fcg_direction forms it), the
// row and column neighbours through the caches, so each plane of z and
// p_old leaves HBM about once.  Per-cell arithmetic as k_fcg_direction's;
// the partial sums add a column's cells first (partial[b], partial[nb + b],
// b the linear block index, nb the block count).
constexpr int kZmX = 64, kZmY = 8;
template <typename T, bool VAR, bool SUMS>
__global__ void __launch_bounds__(kZmX *kZmY)
    k_fcg_direction_zm(DLevel<double> L, const T *__restrict__ z, const double *__restrict__ p_old,
                       double *__restrict__ p, double *__restrict__ q, const double *__restrict__ r,
                       const double *__restrict__ scal, int first, double *__restrict__ partial, int zchunk) {
  const double pq_old = first ? 0.0 : scal[S_PQ];
  const double beta = pq_old != 0.0 ? -scal[S_ZQ] / pq_old : 0.0;  // 0 only after exact convergence
  const int i = blockIdx.x * kZmX + (int)threadIdx.x % kZmX, j = blockIdx.y * kZmY + (int)threadIdx.x / kZmX;
  const int k0 = blockIdx.z * zchunk, k1 = min(k0 + zchunk, L.nz);
  const long sy = L.nx, sz = (long)L.nx * L.ny;
  auto P = [&](long g) { return double(z[g]) + (first ? 0.0 : beta * p_old[g]); };
  double vpq = 0.0, vpr = 0.0, s3[3] = {0.0, 0.0, 0.0};
  if (i < L.nx && j < L.ny && k0 < k1) {
    long c = i + sy * j + sz * k0;
    double pb = k0 > 0 ? P(c - sz) : 0.0, pc = P(c);
    double czb = 0.0;
    if constexpr (VAR) czb = k0 > 0 ? L.cz[c - sz] : 0.0;
    for (int k = k0; k < k1; ++k, c += sz) {
      const double pf = k < L.nz - 1 ? P(c + sz) : 0.0;
      double off = 0.0, diag;
      if constexpr (VAR) {
        const double cR = L.cx[c], cL = i > 0 ? L.cx[c - 1] : 0.0;
        const double cU = L.cy[c], cD = j > 0 ? L.cy[c - sy] : 0.0;
        if (i > 0) off += cL * P(c - 1);
        if (i < L.nx - 1) off += cR * P(c + 1);
        if (j > 0) off += cD * P(c - sy);
        if (j < L.ny - 1) off += cU * P(c + sy);
        diag = -(cR + cL + cU + cD);
        const double cF = __builtin_nontemporal_load(L.cz + c), cB = k > 0 ? czb : 0.0;
        if (k > 0) off += cB * pb;
        if (k < L.nz - 1) off += cF * pf;
        diag -= (cF + cB);
        czb = cF;
      } else {
        double cc = 0.0;
        if (i > 0) { off += L.ax * P(c - 1); cc += L.ax; }
        if (i < L.nx - 1) { off += L.ax * P(c + 1); cc += L.ax; }
        if (j > 0) { off += L.ay * P(c - sy); cc += L.ay; }
        if (j < L.ny - 1) { off += L.ay * P(c + sy); cc += L.ay; }
        if (k > 0) { off += L.az * pb; cc += L.az; }
        if (k < L.nz - 1) { off += L.az * pf; cc += L.az; }
        diag = L.dirichlet ? L.dconst : -cc;
      }
      const double qc = off + diag * pc;
      const double rc = __builtin_nontemporal_load(r + c);
      p[c] = pc;
      q[c] = qc;
      vpq += pc * qc;
      vpr += pc * rc;
      if constexpr (SUMS) {
        s3[0] += pc;
        s3[1] += rc;
        s3[2] += qc;
      }
      pb = pc;
      pc = pf;
    }
  }
  vpq = block_sum(vpq);
  vpr = block_sum(vpr);
  if constexpr (SUMS) block_sum_n<3>(s3);
  if (threadIdx.x == 0) {
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    partial[b] = vpq;
    partial[nb + b] = vpr;
    if constexpr (SUMS) {
      partial[2 * nb + b] = s3[0];
      partial[3 * nb + b] = s3[1];
      partial[4 * nb + b] = s3[2];
    }
  }
}

// 3-D k_residual marching in z (the layout of k_fcg_direction_zm): the
// plane neighbours of x and the face below from a register window, the row
// and column neighbours through the caches; each cell's residual rounds as
// k_residual's (the same operations as stencil(), missing neighbours with a
// zero coefficient), the partial sums add a column's cells first.  The
// right-hand side and the z faces, each read once, are non-temporal loads
// (256^3 FP64 pass 146.6 -> 144.0 us; the x and y faces, which the row and
// column neighbours read again, gain nothing from it), as are the arrays
// that k_rr3d_zm and k_fcg_direction_zm read once (rhs; r and the z faces).
template <bool VAR>
__global__ void __launch_bounds__(kZmX *kZmY)
    k_residual_zm(DLevel<double> L, int store, double *__restrict__ partial, const double *__restrict__ flag,
                  int zchunk) {
  if (flag && *flag != 0.0) return;
  const int i = blockIdx.x * kZmX + (int)threadIdx.x % kZmX, j = blockIdx.y * kZmY + (int)threadIdx.x / kZmX;
  const int k0 = blockIdx.z * zchunk, k1 = min(k0 + zchunk, L.nz);
  const long sy = L.nx, sz = (long)L.nx * L.ny;
  const double *x = L.sol;
  double acc = 0.0;
  if (i < L.nx && j < L.ny && k0 < k1) {
    const bool hl = i > 0, hr = i < L.nx - 1, hd = j > 0, hu = j < L.ny - 1;
    long c = i + sy * j + sz * k0;
    double xb = k0 > 0 ? x[c - sz] : 0.0, xc = x[c];
    double czb = 0.0;
    if constexpr (VAR) czb = k0 > 0 ? L.cz[c - sz] : 0.0;
    for (int k = k0; k < k1; ++k, c += sz) {
      const bool hb = k > 0, hf = k < L.nz - 1;
      const double xf = hf ? x[c + sz] : 0.0;
      const long cl = hl ? c - 1 : c, cr = hr ? c + 1 : c, cd = hd ? c - sy : c, cu = hu ? c + sy : c;
      const double xl = x[cl], xr = x[cr], xd = x[cd], xu = x[cu];
      double off = 0.0, diag;
      if constexpr (VAR) {
        const double cR = L.cx[c], cLv = L.cx[cl], cU = L.cy[c], cDv = L.cy[cd];
        const double cL = hl ? cLv : 0.0, cD = hd ? cDv : 0.0;
        off += cL * xl;
        off += (hr ? cR : 0.0) * xr;
        off += cD * xd;
        off += (hu ? cU : 0.0) * xu;
        diag = -(cR + cL + cU + cD);
        const double cF = __builtin_nontemporal_load(L.cz + c), cB = hb ? czb : 0.0;
        off += cB * xb;
        off += (hf ? cF : 0.0) * xf;
        diag -= (cF + cB);
        czb = cF;
      } else {
        const double wx = L.rbwall ? L.wx : 0.0, wy = L.rbwall ? L.wy : 0.0, wz = L.rbwall ? L.wz : 0.0;
        double cc = 0.0;
        off += (hl ? L.ax : 0.0) * xl;
        cc += hl ? L.ax : wx;
        off += (hr ? L.ax : 0.0) * xr;
        cc += hr ? L.ax : wx;
        off += (hd ? L.ay : 0.0) * xd;
        cc += hd ? L.ay : wy;
        off += (hu ? L.ay : 0.0) * xu;
        cc += hu ? L.ay : wy;
        off += (hb ? L.az : 0.0) * xb;
        cc += hb ? L.az : wz;
        off += (hf ? L.az : 0.0) * xf;
        cc += hf ? L.az : wz;
        diag = (L.dirichlet && !L.rbwall) ? L.dconst : -cc;
      }
      const double rc = __builtin_nontemporal_load(L.rhs + c) - (off + diag * xc);
      if (store) L.res[c] = rc;
      acc += rc * rc;
      xb = xc;
      xc = xf;
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = acc;
}

// 3-D residual + child-average restriction on the A(rho) level 0, marching
// in z (the layout of k_residual_zm): thread (lx, ly) owns fine column (i, j)
// and takes two fine planes per coarse plane, x and rho of the plane below
// and above in a register window, row and column neighbours through the
// caches.  Each residual rounds as k_rr3d_pairs' (rho_face, z3_stencil);
// the eight children are added as: the two planes in the thread, the x pair
// by a lane exchange, the y pair (two rows of the block) through LDS.
// zchunk is even; nx, ny, nz are even (factor-2 coarsening).
template <typename T>
__global__ void __launch_bounds__(kZmX *kZmY)
    k_rr3d_zm(DLevel<T> F, DLevel<T> G, const T *__restrict__ rho, int zchunk) {
  __shared__ T part[kZmY][kZmX];
  const int lx = (int)threadIdx.x % kZmX, ly = (int)threadIdx.x / kZmX;
  const int i = blockIdx.x * kZmX + lx, j = blockIdx.y * kZmY + ly;
  const int k0 = blockIdx.z * zchunk, k1 = min(k0 + zchunk, F.nz);
  const bool in = i < F.nx && j < F.ny;
  const long sy = F.nx, sz = (long)F.nx * F.ny;
  const T inf = T(INFINITY);
  const T *x = F.sol;
  long c = in ? i + sy * j + sz * k0 : 0;
  T xb = T(0), xc = T(0), rb = inf, rcv = inf;
  if (in) {
    xb = k0 > 0 ? x[c - sz] : T(0);
    xc = x[c];
    rb = k0 > 0 ? rho[c - sz] : inf;
    rcv = rho[c];
  }
  for (int k = k0; k < k1; k += 2) {
    T s = T(0);
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int kf = k + kk;
      if (in) {
        const bool hf = kf < F.nz - 1;
        const T xf = hf ? x[c + sz] : T(0), rf = hf ? rho[c + sz] : inf;
        const T cL = rho_face<T>(i > 0 ? rho[c - 1] : inf, rcv, F.ax);
        const T cR = rho_face<T>(rcv, i < F.nx - 1 ? rho[c + 1] : inf, F.ax);
        const T cD = rho_face<T>(j > 0 ? rho[c - sy] : inf, rcv, F.ay);
        const T cU = rho_face<T>(rcv, j < F.ny - 1 ? rho[c + sy] : inf, F.ay);
        const T cB = rho_face<T>(rb, rcv, F.az);
        const T cF = rho_face<T>(rcv, rf, F.az);
        T off, diag;
        z3_stencil<T, 2>(F, i, j, kf, i > 0 ? x[c - 1] : T(0), i < F.nx - 1 ? x[c + 1] : T(0),
                         j > 0 ? x[c - sy] : T(0), j < F.ny - 1 ? x[c + sy] : T(0), xb, xf, cL, cR, cD, cU, cB,
                         cF, off, diag);
        s += __builtin_nontemporal_load(F.rhs + c) - (off + diag * xc);
        xb = xc;
        xc = xf;
        rb = rcv;
        rcv = rf;
        c += sz;
      }
    }
    s += __shfl_xor(s, 1, kWave);
    part[ly][lx] = s;
    __syncthreads();
    if (in && (ly & 1) == 0 && (lx & 1) == 0) {
      const long C = (i >> 1) + (long)G.nx * ((j >> 1) + (long)G.ny * (k >> 1));
      G.rhs[C] = (part[ly][lx] + part[ly + 1][lx]) / T(8);
      G.sol[C] = T(0);
    }
    __syncthreads();
  }
}

// 3-D factor-2 prolong-add marching in z: thread (lx, ly) owns the fine
// x pair 2m, 2m+1 (m = bx kZmX + lx) of fine row j for an even chunk of
// planes, with the three coarse planes the next two fine planes need in a
// register window (fine planes 2K', 2K'+1 read coarse planes K'-1, K', K'+1),
// so each coarse plane is loaded once per column instead of four times.
// Same weights, ghost rule and operation order as k_prolong3_r2.
template <typename T>
__global__ void __launch_bounds__(kZmX *kZmY) k_prolong3_zm(DLevel<T> G, DLevel<T> F, int dir, int zchunk) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const int lx = (int)threadIdx.x % kZmX, ly = (int)threadIdx.x / kZmX;
  const int m = blockIdx.x * kZmX + lx, j = blockIdx.y * kZmY + ly;
  const int k0 = blockIdx.z * zchunk, k1 = min(k0 + zchunk, F.nz);
  if (m >= G.nx || j >= F.ny) return;  // no barriers below
  const int J0 = (j - 1) >> 1;
  const T ty = (j & 1) ? T(0.25) : T(0.75);
  auto load = [&](T (&a)[2][3], int K) {  // coarse plane K, ghost rule of coarse_at
    const bool kout = K < 0 || K >= G.nz;
    const int Kc = min(max(K, 0), G.nz - 1);
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int J = J0 + jj;
      const bool rout = kout || J < 0 || J >= G.ny;
      const T *row = G.sol + (long)G.nx * (min(max(J, 0), G.ny - 1) + (long)G.ny * Kc);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int I = m - 1 + d;
        const bool out = rout || I < 0 || I >= G.nx;
        a[jj][d] = (out && dir) ? T(0) : row[min(max(I, 0), G.nx - 1)];
      }
    }
  };
  T pa[2][3], pb[2][3], pc[2][3];
  int Kp = k0 >> 1;
  load(pa, Kp - 1);
  load(pb, Kp);
  load(pc, Kp + 1);
  T *fine = F.sol + 2 * m + (long)F.nx * (j + (long)F.ny * k0);
  const long sz = (long)F.nx * F.ny;
  for (int k = k0; k < k1; k += 2, ++Kp) {
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const T tz = kk ? T(0.25) : T(0.75);
      const T(&A)[2][3] = kk ? pb : pa;
      const T(&B)[2][3] = kk ? pc : pb;
      auto interp = [&](T tx, int o) {
        const T v00 = (T(1) - tx) * A[0][o] + tx * A[0][o + 1];
        const T v10 = (T(1) - tx) * A[1][o] + tx * A[1][o + 1];
        const T v0 = (T(1) - ty) * v00 + ty * v10;
        const T v01 = (T(1) - tx) * B[0][o] + tx * B[0][o + 1];
        const T v11 = (T(1) - tx) * B[1][o] + tx * B[1][o + 1];
        const T v1 = (T(1) - ty) * v01 + ty * v11;
        return (T(1) - tz) * v0 + tz * v1;
      };
      T2 *p = reinterpret_cast<T2 *>(fine + (kk ? sz : 0));
      T2 f = *p;
      f.x += interp(T(0.75), 0);
      f.y += interp(T(0.25), 1);
      *p = f;
    }
    fine += 2 * sz;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        pa[jj][d] = pb[jj][d];
        pb[jj][d] = pc[jj][d];
      }
    if (k + 2 < k1) load(pc, Kp + 2);
  }
}

// x += alpha p, r -= alpha q (alpha = p.r / p.q); r32 = float(r) when given;
// partial sums of r.r.  p.q and p.r from scal, or, with dir_nb > 0, reduced
// here from k_fcg_direction's partials (p.q at [0, dir_nb), p.r after them),
// in one order for every workgroup, block 0 storing them in scal for the
// next beta.
// partial must not overlap dir_part.
__global__ void __launch_bounds__(kFcgThreads)
    k_fcg_update(long n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                 const double *__restrict__ q, double *__restrict__ scal, float *__restrict__ r32,
                 double *__restrict__ partial, const double *__restrict__ dir_part, int dir_nb, double corr_n) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double pq, pr;
  if (dir_nb > 0) {
    fcg_dir_sums(dir_part, dir_nb, corr_n, pq, pr);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      scal[S_PQ] = pq;
      scal[S_PR] = pr;
    }
  } else {
    pq = scal[S_PQ];
    pr = scal[S_PR];
  }
  double v = 0.0;
  if (c < n) {
    const double alpha = pq != 0.0 ? pr / pq : 0.0;  // 0 only after exact convergence
    x[c] += alpha * p[c];
    const double rc = r[c] - alpha * q[c];
    r[c] = rc;
    if (r32) r32[c] = float(rc);
    v = rc * rc;
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kBlock) k_copy_cast(const double *__restrict__ a, double *__restrict__ b,
                                                      float *__restrict__ b32, long n) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    if (b) b[c] = a[c];
    if (b32) b32[c] = float(a[c]);
  }
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

// ------------------------------------------------------ learned guess
// network features on a level of block size bs (W x H x D cells, each the
// mean of bs^ndim level-0 cells): the block mean of b scaled by 1/sigma_b
// and the block mean of ln(rho) / ln(1000); optionally the block mean of the
// level-0 solution scaled by 1/sigma_b (training target)
__global__ void __launch_bounds__(kBlock) k_learned_features(const double *__restrict__ rhs0,
                                                             const double *__restrict__ rho0,
                                                             const double *__restrict__ sol0, int nx0, int ny0,
                                                             int bs, int bsz, int W, int H, int D, double inv_sigma,
                                                             float *__restrict__ feat, float *__restrict__ tgt) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)W * H * D) return;
  const int x = (int)(p % W), y = (int)((p / W) % H), z = (int)(p / ((long)W * H));
  double sb = 0.0, sr = 0.0, sp = 0.0;
  for (int e = 0; e < bsz; ++e)
    for (int b = 0; b < bs; ++b)
      for (int a = 0; a < bs; ++a) {
        const long c = ((long)(z * bsz + e) * ny0 + (y * bs + b)) * nx0 + (x * bs + a);
        sb += rhs0[c];
        if (rho0) sr += log(rho0[c]);
        if (sol0) sp += sol0[c];
      }
  const double inv = 1.0 / ((double)bs * bs * bsz);
  feat[2 * p] = (float)(sb * inv * inv_sigma);
  feat[2 * p + 1] = (float)(sr * inv / 6.907755278982137);  // ln(1000)
  if (tgt) tgt[p] = (float)(sp * inv * inv_sigma);
}

// faces of a level scaled by its h^2 (face coefficient x h^2 = 1 / mean rho
// of the face), network layout face[p] = (x+1 face, y+1 face (, z+1 face)),
// zero on walls
__global__ void __launch_bounds__(kBlock) k_learned_faces(const double *__restrict__ cx, const double *__restrict__ cy,
                                                          const double *__restrict__ cz, int W, int H, int D, int ndim,
                                                          double hx2, double hy2, double hz2,
                                                          float *__restrict__ face) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)W * H * D) return;
  const int x = (int)(p % W), y = (int)((p / W) % H), z = (int)(p / ((long)W * H));
  face[ndim * p] = x + 1 < W ? (float)(cx[p] * hx2) : 0.f;
  face[ndim * p + 1] = y + 1 < H ? (float)(cy[p] * hy2) : 0.f;
  if (ndim == 3) face[3 * p + 2] = z + 1 < D ? (float)(cz[p] * hz2) : 0.f;
}

__global__ void __launch_bounds__(kBlock) k_net_to_level(const float *__restrict__ out, double *__restrict__ sol,
                                                         long n, double scale) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) sol[c] = (double)out[c] * scale;
}

// four inner products of a guess p with q = A p: p.b, p.q, b.q, q.q
// (partial[k*nb + block]); they give the energy-optimal scale of p and the
// residual norm of the scaled guess without another pass
template <int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock) k_guess_dots(DLevel<double> L, const double *__restrict__ p,
                                                       double *__restrict__ partial, int nb) {
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  if (c < L.n) {
    int i, j, k;
    cell_ijk(c, L.nx, L.ny, i, j, k);
    double off, diag;
    stencil<double, DIM, VAR>(L, p, i, j, k, c, off, diag);
    const double q = off + diag * p[c], b = L.rhs[c];
    v0 = p[c] * b;
    v1 = p[c] * q;
    v2 = b * q;
    v3 = q * q;
  }
  v0 = block_sum(v0);
  v1 = block_sum(v1);
  v2 = block_sum(v2);
  v3 = block_sum(v3);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = v0;
    partial[nb + blockIdx.x] = v1;
    partial[2 * nb + blockIdx.x] = v2;
    partial[3 * nb + blockIdx.x] = v3;
  }
}

// x.b, x.r and g.r of an FCG iterate x, its residual r and a learned guess g
// (partial[k * nb + block], k = 0, 1, 2)
__global__ void __launch_bounds__(kBlock) k_combine_dots(const double *__restrict__ x, const double *__restrict__ r,
                                                         const double *__restrict__ b, const double *__restrict__ g,
                                                         long n, double *__restrict__ partial, int nb) {
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (c < n) {
    const double xc = x[c], rc = r[c];
    v0 = xc * b[c];
    v1 = xc * rc;
    v2 = g[c] * rc;
  }
  v0 = block_sum(v0);
  v1 = block_sum(v1);
  v2 = block_sum(v2);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = v0;
    partial[nb + blockIdx.x] = v1;
    partial[2 * nb + blockIdx.x] = v2;
  }
}

// y = a x + c y
__global__ void __launch_bounds__(kBlock) k_axpby(const double *__restrict__ x, double *__restrict__ y, long n,
                                                  double a, double c) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + c * y[i];
}

// least gain of the energy-optimal combination of an FCG iterate and the
// learned guess, as a fraction of |J(x_k)| (learned_combine)
constexpr double kCombineGain = 0.05;

// x += y
__global__ void __launch_bounds__(kBlock) k_add(double *__restrict__ x, const double *__restrict__ y, long n) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) x[c] += y[c];
}

__global__ void __launch_bounds__(kBlock) k_scale(double *__restrict__ x, long n, double a) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) x[c] *= a;
}

// ============================================================ host side
template <typename T>
struct Level {
  int nx = 0, ny = 0, nz = 0, deg = 1;
  long n = 0;
  double dx = 1, dy = 1, dz = 1;
  // `sol` is the current iterate; the fused sweep writes `sol_alt` and the
  // two pointers are swapped afterwards
  T *sol = nullptr, *sol_alt = nullptr, *rhs = nullptr, *res = nullptr, *buf = nullptr;
  T *rhs_alt = nullptr;  // FP32 level 0: next right-hand side written by the fused outer step
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

// Tile shapes of the LDS-tiled 2-D kernels, largest first.  A level uses the
// largest shape that still gives every CU two workgroups (small levels are
// latency bound: parallelism beats the halo overhead of small tiles).
template <typename T>
struct TileCfg;
template <>
struct TileCfg<double> {  // largest: 4 x 74 x 22 x 8 B + coarse tile = 57 KiB of LDS
  static constexpr int TX[4] = {64, 32, 16, 8}, TY[4] = {12, 16, 16, 8};
};
template <>
struct TileCfg<float> {  // largest: 4 x 74 x 42 x 4 B + coarse tile = 54 KiB of LDS
  static constexpr int TX[4] = {64, 32, 16, 8}, TY[4] = {32, 16, 16, 8};
};

}  // namespace scihpc

using namespace scihpc;

struct scihpc_mg {
  int ndim = 2, bc = 0, L = 0;
  bool variable = false;
  std::vector<Level<double>> d;
  std::vector<Level<float>> f;  // FP32 hierarchy for the mixed path (lazy)
  std::vector<int> nocompat;
  double *Minv = nullptr;       // dense inverse of the (pinned) coarsest red-black operator
  double *Minv_ref = nullptr;   // same for the reference operator (differs for Dirichlet)
  double *partial = nullptr;    // per-block partial sums
  double *scal = nullptr;       // device scalars (S_*)
  double *host_scal = nullptr;  // pinned mirror
  double *p_alt = nullptr;      // FP64 ping-pong buffer for the outer loop
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  scihpc::LearnedHook learned;
  long block_max = 1L << 30;  // optional cap on the one-workgroup cycle's top level
  long lds_max = 65536;       // LDS bytes a workgroup may use (queried at create)
  int num_cu = 256;           // compute units (queried at create)
  bool tiled = true;          // LDS-tiled 2-D fine-level kernels (SCIHPC_MG_TILED=0 disables)
  bool tiled3d = true;        // z-streaming 3-D sweeps (SCIHPC_MG_TILED3D=0 disables)
  bool fused_update = true;   // mixed 2-D FCG: update inside the level-0 pre-smoothing
                              // (SCIHPC_MG_FUSED_UPDATE=0 disables)
  int z3_chunk = 0;           // z planes per 3-D workgroup (0: chosen per level)
  float *rho0f = nullptr;     // FP32 copy of rho0 for the mixed V-cycle
  void *blk_dev[2] = {nullptr, nullptr};   // device copies of BlockLevels<double>/<float>
  std::vector<char> blk_host[2];           // last uploaded contents
  double *host_part = nullptr;             // pinned copy of the partial sums
  double *rho0 = nullptr;                  // level-0 density (MG layout) of the variable operator
  float *lfeat = nullptr, *lout = nullptr; // learned-guess features and network output

  // ------------------------------------------------------------ views
  // rb: view of the red-black solver's operator (Dirichlet coarse levels
  // get the boundary-consistent wall term); false: the reference operator
  template <typename T>
  DLevel<T> view(const Level<T> &l, bool rb = true) const {
    DLevel<T> v;
    const auto &l0 = d[0];
    v.rbwall = rb && bc == SCIHPC_MG_DIRICHLET && !variable && l.deg > 1;
    auto wall = [](double H, double h0) { return 2.0 / (H * H * (1.0 + h0 / H)); };
    v.wx = T(wall(l.dx, l0.dx));
    v.wy = T(wall(l.dy, l0.dy));
    v.wz = T(ndim == 3 ? wall(l.dz, l0.dz) : 0.0);
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
    else if (r == 2 && fine.n >= (1L << 18) && ceil_div(fine.ny, kZmY) <= 65535) {
      // z-march: about 2048 blocks, chunks of at least 8 (even) planes
      const long bxy = (long)ceil_div(coarse.nx, kZmX) * ceil_div(fine.ny, kZmY);
      const long slabs = std::max(1L, 2048L / bxy);
      const int chunk = std::max(8, 2 * (int)ceil_div(fine.nz / 2, slabs));
      const dim3 grid(ceil_div(coarse.nx, kZmX), ceil_div(fine.ny, kZmY), ceil_div(fine.nz, chunk));
      k_prolong3_zm<T><<<grid, kZmX * kZmY, 0, stream>>>(view(coarse), view(fine), dir, chunk);
    } else if (r == 2) {
      SCIHPC_REQUIRE(fine.nz <= 65535, "3-D level too tall for the prolongation grid");
      const dim3 grid(ceil_div(coarse.nx, kP3X), ceil_div(fine.ny, kP3Y), fine.nz);
      k_prolong3_r2<T><<<grid, dim3(kP3X, kP3Y), 0, stream>>>(view(coarse), view(fine), dir);
    } else
      k_prolong_add<T, 3><<<ceil_div(fine.n, kBlock), kBlock, 0, stream>>>(view(coarse), view(fine), r, dir);
    SCIHPC_CHECK_LAUNCH();
  }
  template <typename T>
  void coarse_solve(const Level<T> &c) {
    k_coarse_solve<T><<<1, kBlock, 0, stream>>>(view(c), Minv, bc == SCIHPC_MG_NEUMANN);
    SCIHPC_CHECK_LAUNCH();
  }

  // ------------------------------------------------ tiled 2-D kernels
  template <typename T>
  bool fused_level(const std::vector<Level<T>> &lv, int l) const {
    return tiled && ndim == 2 && l + 1 < L && lv[l + 1].deg / lv[l].deg == 2;
  }
  // tile shape index for a level (see TileCfg) and its tile count
  template <typename T>
  int tile_kind(const Level<T> &l, int level) const {
    // the largest tile that still gives two workgroups per CU; 8 x 8 on the
    // small latency-bound coarse levels (shorter chains per workgroup).
    // Level 0 keeps at most one tile per 256 cells: its per-tile partial
    // sums share the partial buffer with the FCG kernels'.
    for (int t = 0; t < 3; ++t)
      if ((long)ceil_div(l.nx, TileCfg<T>::TX[t]) * ceil_div(l.ny, TileCfg<T>::TY[t]) >=
          (t == 2 && level > 0 ? 4L : 2L) * num_cu)
        return t;
    return level == 0 ? 2 : 3;
  }
  template <typename T>
  unsigned tiles2d(const Level<T> &l, int level) const {
    const int t = tile_kind(l, level);
    return ceil_div(l.nx, TileCfg<T>::TX[t]) * ceil_div(l.ny, TileCfg<T>::TY[t]);
  }

  template <typename T, bool VAR, int K>
  void pre2d_k(std::vector<Level<T>> &lv, int l, int nu1, int zero_guess) {
    constexpr int TX = TileCfg<T>::TX[K], TY = TileCfg<T>::TY[K];
    auto &a = lv[l];
    // 8 x 8 tiles (18 x 18 staged cells) run with four waves
    constexpr int threads = TX * TY <= 64 ? kFusedThreads / 2 : kFusedThreads;
    k_pre2d<T, VAR, TX, TY><<<tiles2d(a, l), threads, 0, stream>>>(view(a), view(lv[l + 1]), a.sol_alt, nu1,
                                                                        zero_guess);
    SCIHPC_CHECK_LAUNCH();
    std::swap(a.sol, a.sol_alt);
  }
  template <typename T>
  void pre2d(std::vector<Level<T>> &lv, int l, int nu1, int zero_guess) {
    const int k = tile_kind(lv[l], l);
    if (variable) {
      if (k == 0) pre2d_k<T, true, 0>(lv, l, nu1, zero_guess);
      else if (k == 1) pre2d_k<T, true, 1>(lv, l, nu1, zero_guess);
      else if (k == 2) pre2d_k<T, true, 2>(lv, l, nu1, zero_guess);
      else pre2d_k<T, true, 3>(lv, l, nu1, zero_guess);
    } else {
      if (k == 0) pre2d_k<T, false, 0>(lv, l, nu1, zero_guess);
      else if (k == 1) pre2d_k<T, false, 1>(lv, l, nu1, zero_guess);
      else if (k == 2) pre2d_k<T, false, 2>(lv, l, nu1, zero_guess);
      else pre2d_k<T, false, 3>(lv, l, nu1, zero_guess);
    }
  }

  // measurement hook: HIP events around the launches of one kernel inside
  // solves (probe_which 6: k_pre2d_fcg), at most probe_ev.size() / 2
  int probe_which = -1, probe_used = 0;
  std::vector<hipEvent_t> probe_ev;
  unsigned long long *probe_clk = nullptr;  // per launch: start and end of every workgroup
  long probe_stride = 0;                     // clock slots per launch
  bool probe_on(int which) const { return probe_which == which && 2 * probe_used < (int)probe_ev.size(); }

  // the FCG update fused into the FP32 level-0 pre-smoothing (k_pre2d_fcg)
  template <bool VAR, int K>
  void pre2d_fcg_k(int nu1, const FcgUpdate &u) {
    constexpr int TX = TileCfg<float>::TX[K], TY = TileCfg<float>::TY[K];
    auto &a = f[0];
    const bool probe = probe_on(6);
    FcgUpdate uc = u;
    if (probe) {
      SCIHPC_REQUIRE(2L * tiles2d(a, 0) <= probe_stride, "internal: probe clock slots");
      uc.clk = probe_clk + probe_stride * probe_used;
      SCIHPC_HIP(hipEventRecord(probe_ev[2 * probe_used], stream));
    }
    k_pre2d_fcg<VAR, TX, TY><<<tiles2d(a, 0), kFusedThreads, 0, stream>>>(view(a), view(f[1]), a.sol_alt, nu1, uc);
    SCIHPC_CHECK_LAUNCH();
    if (probe) SCIHPC_HIP(hipEventRecord(probe_ev[2 * probe_used++ + 1], stream));
    std::swap(a.sol, a.sol_alt);
  }
  void pre2d_fcg(int nu1, const FcgUpdate &u) {
    const int k = tile_kind(f[0], 0);
    if (variable) {
      if (k == 0) pre2d_fcg_k<true, 0>(nu1, u);
      else if (k == 1) pre2d_fcg_k<true, 1>(nu1, u);
      else pre2d_fcg_k<true, 2>(nu1, u);  // level 0: never the 8 x 8 kind
    } else {
      if (k == 0) pre2d_fcg_k<false, 0>(nu1, u);
      else if (k == 1) pre2d_fcg_k<false, 1>(nu1, u);
      else pre2d_fcg_k<false, 2>(nu1, u);  // level 0: never the 8 x 8 kind
    }
  }

  template <typename T, bool VAR, int K, int MODE>
  void post2d_k(std::vector<Level<T>> &lv, int l, int nu2, const OuterArgs &oa) {
    constexpr int TX = TileCfg<T>::TX[K], TY = TileCfg<T>::TY[K];
    auto &a = lv[l];
    const int dir = bc == SCIHPC_MG_DIRICHLET;
    constexpr int threads = TX * TY <= 64 ? kFusedThreads / 2 : kFusedThreads;
    k_post2d<T, VAR, TX, TY, MODE><<<tiles2d(a, l), threads, 0, stream>>>(view(lv[l + 1]), view(a), a.sol_alt,
                                                                               nu2, dir, partial, oa);
    SCIHPC_CHECK_LAUNCH();
    if (MODE != 2) std::swap(a.sol, a.sol_alt);
  }
  template <typename T, int MODE>
  void post2d_mode(std::vector<Level<T>> &lv, int l, int nu2, const OuterArgs &oa) {
    const int k = tile_kind(lv[l], l);
    if (variable) {
      if (k == 0) post2d_k<T, true, 0, MODE>(lv, l, nu2, oa);
      else if (k == 1) post2d_k<T, true, 1, MODE>(lv, l, nu2, oa);
      else if (k == 2) post2d_k<T, true, 2, MODE>(lv, l, nu2, oa);
      else post2d_k<T, true, 3, MODE>(lv, l, nu2, oa);
    } else {
      if (k == 0) post2d_k<T, false, 0, MODE>(lv, l, nu2, oa);
      else if (k == 1) post2d_k<T, false, 1, MODE>(lv, l, nu2, oa);
      else if (k == 2) post2d_k<T, false, 2, MODE>(lv, l, nu2, oa);
      else post2d_k<T, false, 3, MODE>(lv, l, nu2, oa);
    }
  }

  // ------------------------------------------- z-streaming 3-D kernels
  template <typename T>
  bool z3_level(const std::vector<Level<T>> &lv, int l) const {
    const auto &a = lv[l];
    return tiled3d && ndim == 3 && a.nx >= 16 && a.ny >= 8 && a.nz >= 8;
  }
  // coefficient source of a level: 2 density (level 0 of A(rho)), 1 faces, 0 constant
  template <typename T>
  int z3_coef(int l) const {
    if (!variable) return 0;
    return (l == 0 && rho0 && (std::is_same<T, double>::value || rho0f)) ? 2 : 1;
  }
  template <typename T>
  const T *z3_rho() const {
    if constexpr (std::is_same<T, double>::value) return rho0;
    else return rho0f;
  }
  int z3_chunk_for(long ntiles, int nz) const {
    if (z3_chunk > 0) return std::min(z3_chunk, nz);
    long chunks = std::max(1L, (4L * num_cu + ntiles - 1) / ntiles);
    chunks = std::min<long>(chunks, std::max(1, nz / 8));
    return (int)((nz + chunks - 1) / chunks);
  }
  template <typename T, int COEF, int TX, int TY>
  void sweep3_k(Level<T> &a, const T *rho, int first, int zero_guess) {
    const long ntiles = (long)ceil_div(a.nx, TX) * ceil_div(a.ny, TY);
    const int zc = z3_chunk_for(ntiles, a.nz);
    const unsigned grid = (unsigned)(ntiles * ceil_div(a.nz, zc));
    k_sweep3d<T, COEF, TX, TY><<<grid, kZ3Threads, 0, stream>>>(view(a), a.sol, a.sol_alt, rho, T(a.dx), T(a.dy),
                                                                T(a.dz), first, zero_guess, zc);
    SCIHPC_CHECK_LAUNCH();
    std::swap(a.sol, a.sol_alt);
  }
  template <typename T>
  void sweep3(std::vector<Level<T>> &lv, int l, int first, int zero_guess) {
    auto &a = lv[l];
    const int coef = z3_coef<T>(l);
    constexpr bool dbl = std::is_same<T, double>::value;
    if (coef == 2) sweep3_k<T, 2, 32, dbl ? 8 : 16>(a, z3_rho<T>(), first, zero_guess);
    else if (coef == 1) sweep3_k<T, 1, dbl ? 16 : 32, dbl ? 8 : 16>(a, nullptr, first, zero_guess);
    else sweep3_k<T, 0, 32, dbl ? 8 : 16>(a, nullptr, first, zero_guess);
  }
  template <typename T>
  void rr3(std::vector<Level<T>> &lv, int l) {
    auto &a = lv[l], &b = lv[l + 1];
    const int coef = z3_coef<T>(l);
    if (coef == 2 && a.n >= (1L << 18) && ceil_div(a.ny, kZmY) <= 65535) {
      // z-march on the A(rho) level: about 8192 blocks, even plane chunks
      const long bxy = (long)ceil_div(a.nx, kZmX) * ceil_div(a.ny, kZmY);
      const long slabs = std::max(1L, 8192L / bxy);
      const int chunk = 2 * (int)ceil_div(a.nz / 2, slabs);
      const dim3 grid(ceil_div(a.nx, kZmX), ceil_div(a.ny, kZmY), ceil_div(a.nz, chunk));
      k_rr3d_zm<T><<<grid, kZmX * kZmY, 0, stream>>>(view(a), view(b), z3_rho<T>(), chunk);
      SCIHPC_CHECK_LAUNCH();
      return;
    }
    SCIHPC_REQUIRE(b.ny <= 65535 && b.nz <= 65535, "3-D level too tall for the restriction grid");
    const dim3 grid(ceil_div(b.nx, kRR3X), b.ny, b.nz), block(kRR3X, 2, 2);
    if (coef == 2) k_rr3d_pairs<T, 2><<<grid, block, 0, stream>>>(view(a), view(b), z3_rho<T>());
    else if (coef == 1) k_rr3d_pairs<T, 1><<<grid, block, 0, stream>>>(view(a), view(b), nullptr);
    else k_rr3d_pairs<T, 0><<<grid, block, 0, stream>>>(view(a), view(b), nullptr);
    SCIHPC_CHECK_LAUNCH();
  }

  // generic full sweep (3-D levels, ratio-3 levels): two colour kernels
  template <typename T>
  void sweep(Level<T> &l, int first, int zero_guess) {
    gs_color(l, first, zero_guess);
    gs_color(l, 1 - first, 0);
  }

  // first level cycled by the one-workgroup LDS kernel (L when even the
  // coarsest level does not fit the LDS budget)
  template <typename T>
  int block_level(const std::vector<Level<T>> &lv) const {
    const long bpc = (long)sizeof(T) * (2 + (variable ? ndim : 0));
    const long budget = lds_max - 2048;
    long used = 0;
    int lb = L;
    for (int l = L - 1; l >= 0 && l >= L - kMaxBlockLevels; --l) {
      used += lv[l].n * bpc;
      if (used > budget || lv[l].n > block_max) break;
      lb = l;
    }
    return lb;
  }

  template <typename T>
  void block_cycle(std::vector<Level<T>> &lv, int lb, int nu1, int nu2, int zero_guess0, int load_top) {
    BlockLevels<T> B;
    std::memset(&B, 0, sizeof(B));
    B.nlev = L - lb;
    B.arrays = 2 + (variable ? ndim : 0);
    long off = 0;
    for (int i = 0; i < B.nlev; ++i) {
      B.lv[i] = view(lv[lb + i]);
      B.ratio[i] = i + 1 < B.nlev ? lv[lb + i + 1].deg / lv[lb + i].deg : 1;
      B.off[i] = (int)off;
      off += lv[lb + i].n * B.arrays;
    }
    const size_t bytes = sizeof(T) * off;
    const int neu = bc == SCIHPC_MG_NEUMANN, dir = bc == SCIHPC_MG_DIRICHLET;
    // the metadata lives in device memory (a by-value kernel argument indexed
    // by level would be copied to scratch); re-uploaded only when it changes
    const int slot = std::is_same<T, double>::value ? 0 : 1;
    if (!blk_dev[slot]) SCIHPC_HIP(hipMalloc(&blk_dev[slot], sizeof(BlockLevels<double>)));
    if (blk_host[slot].size() != sizeof(B) || std::memcmp(blk_host[slot].data(), &B, sizeof(B)) != 0) {
      blk_host[slot].assign(reinterpret_cast<const char *>(&B), reinterpret_cast<const char *>(&B) + sizeof(B));
      SCIHPC_HIP(hipMemcpyAsync(blk_dev[slot], blk_host[slot].data(), sizeof(B), hipMemcpyHostToDevice, stream));
      SCIHPC_HIP(hipStreamSynchronize(stream));
    }
    const BlockLevels<T> *Bd = static_cast<const BlockLevels<T> *>(blk_dev[slot]);
#define SCIHPC_BLOCK(D, V)                                                                                  \
  do {                                                                                                      \
    if (bytes > 65536)                                                                                      \
      SCIHPC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_block_vcycle<T, D, V>),             \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));              \
    k_block_vcycle<T, D, V><<<1, kBlockThreads, bytes, stream>>>(Bd, nu1, nu2, Minv, neu, dir, zero_guess0, \
                                                                  load_top);                                \
  } while (0)
    if (ndim == 2) {
      if (variable) SCIHPC_BLOCK(2, true);
      else SCIHPC_BLOCK(2, false);
    } else {
      if (variable) SCIHPC_BLOCK(3, true);
      else SCIHPC_BLOCK(3, false);
    }
#undef SCIHPC_BLOCK
    SCIHPC_CHECK_LAUNCH();
  }

  // V(nu1,nu2) on hierarchy `lv`, in place on lv[0].sol.  Fine 2-D levels
  // use the tiled pre/post kernels, the coarse levels one LDS workgroup.
  // post0 selects what the level-0 post-smoothing kernel also does:
  // 0 nothing, 1 FP64 residual norm (partials), 2 mixed outer step (oa).
  // Returns the number of residual partials written (0 when post0 == 0).
  template <typename T>
  // pre0_done: the level-0 pre-smoothing has already run (k_pre2d_fcg)
  int vcycle(std::vector<Level<T>> &lv, int nu1, int nu2, bool zero_guess0, int post0 = 0,
             const OuterArgs &oa = OuterArgs{}, bool pre0_done = false) {
    const int lb = block_level(lv);
    if (nu1 == 0 && zero_guess0) SCIHPC_HIP(hipMemsetAsync(lv[0].sol, 0, sizeof(T) * lv[0].n, stream));
    const bool zg = zero_guess0 && nu1 > 0;
    const bool can_tile = nu1 <= 2 && nu2 <= 2;  // the tiles stage a 5-cell halo
    const int top = std::min(lb, L - 1);  // deepest level reached by the fine loop
    SCIHPC_REQUIRE(!pre0_done || (top >= 1 && can_tile && fused_level(lv, 0)), "internal: level 0 not tiled");
    for (int l = pre0_done ? 1 : 0; l < top; ++l) {
      const int z = (l == 0 && zg) ? 1 : 0;
      if (can_tile && fused_level(lv, l)) {
        pre2d(lv, l, nu1, z);
      } else if (z3_level(lv, l) && lv[l + 1].deg / lv[l].deg == 2) {
        for (int s = 0; s < nu1; ++s) sweep3(lv, l, 0, s == 0 ? z : 0);
        rr3(lv, l);
      } else {
        for (int s = 0; s < nu1; ++s) sweep(lv[l], 0, s == 0 ? z : 0);
        residual_restrict(lv[l], lv[l + 1]);
      }
    }
    if (lb < L) {
      block_cycle(lv, lb, nu1, nu2, (lb == 0 && zg) ? 1 : 0, lb == 0 ? 1 : 0);
    } else {
      coarse_solve(lv[L - 1]);
    }
    int partials = 0;
    bool post0_done = false;
    for (int l = top; l > 0; --l) {
      const int f = l - 1;
      if (can_tile && fused_level(lv, f)) {
        const int mode = f == 0 ? post0 : 0;
        if (mode == 0) post2d_mode<T, 0>(lv, f, nu2, oa);
        else if (mode == 1) post2d_mode<T, 1>(lv, f, nu2, oa);
        else if (mode == 2) post2d_mode<T, 2>(lv, f, nu2, oa);
        else post2d_mode<T, 3>(lv, f, nu2, oa);
        if (f == 0) {
          post0_done = true;
          partials = mode != 0 ? (int)tiles2d(lv[0], 0) : 0;
        }
      } else if (z3_level(lv, f)) {
        prolong_add(lv[l], lv[f]);
        for (int s = 0; s < nu2; ++s) sweep3(lv, f, 1, 0);
      } else {
        prolong_add(lv[l], lv[f]);
        for (int s = 0; s < nu2; ++s) sweep(lv[f], 1, 0);
      }
    }
    if (post0 != 0 && !post0_done) partials = post0_generic(lv, post0, oa);
    return partials;
  }

  // level-0 post work without the tiled kernel (3-D, or one-block cycles):
  // 1 FP64 residual norm, 2 mixed outer step, 3 z.q for the flexible CG;
  // returns the number of per-block partial sums written
  template <typename T>
  int post0_generic(std::vector<Level<T>> &lv, int post0, const OuterArgs &oa) {
    const unsigned nb = ceil_div(lv[0].n, kBlock);
    if (post0 == 3) {
      k_dot_zq<T><<<nb, kBlock, 0, stream>>>(lv[0].sol, oa.q, lv[0].n, partial);
      SCIHPC_CHECK_LAUNCH();
      return (int)nb;
    }
    if constexpr (std::is_same<T, double>::value) {
      return residual_partials(lv[0], false);
    } else {
      // e = lv[0].sol; write p_new and float(r) into oa.r32
#define SCIHPC_OUTER(D, V)                                                                                     \
  k_outer_update<D, V><<<nb, kBlock, 0, stream>>>(oa.P, oa.p_old, lv[0].sol, oa.p_new, oa.r32, partial, 0)
      if (ndim == 2) { if (variable) SCIHPC_OUTER(2, true); else SCIHPC_OUTER(2, false); }
      else { if (variable) SCIHPC_OUTER(3, true); else SCIHPC_OUTER(3, false); }
#undef SCIHPC_OUTER
      SCIHPC_CHECK_LAUNCH();
    }
    return (int)nb;
  }

  // sum of partial[0..np) on the host (the norm the host tests anyway):
  // one D2H copy, fixed summation order
  // partial sums beyond kCompressAbove are first summed in groups of 256 on
  // the device (one wide launch instead of one long single-block loop)
  static constexpr int kCompressAbove = 2048;
  double *partial2 = nullptr;
  int compress(const double *src, int n, double *dst) {
    const int m = (n + kBlock - 1) / kBlock;
    k_compress<<<m, kBlock, 0, stream>>>(src, n, kBlock, dst);
    SCIHPC_CHECK_LAUNCH();
    return m;
  }
  double host_sum(int np) {
    const double *src = partial;
    if (np > kCompressAbove) {
      np = compress(partial, np, partial2);
      src = partial2;
    }
    SCIHPC_HIP(hipMemcpyAsync(host_part, src, sizeof(double) * np, hipMemcpyDeviceToHost, stream));
    SCIHPC_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (int q = 0; q < np; ++q) s += host_part[q];
    return s;
  }

  // sum of partial[0..nb) -> scal[slot]; returns nothing (device only)
  void finalize(int nb, int slot, int op = 0) {
    if (nb > kCompressAbove) {
      const int m = compress(partial, nb, partial2);
      k_finalize<<<1, kBlock, 0, stream>>>(partial2, m, scal, slot, op);
    } else {
      k_finalize<<<1, kBlock, 0, stream>>>(partial, nb, scal, slot, op);
    }
    SCIHPC_CHECK_LAUNCH();
  }
  // scal[s0] = sum partial[0..nb), scal[s1] = sum partial[nb..2nb)
  void finalize2(int nb, int s0, int s1) {
    if (nb > kCompressAbove) {
      const int m = compress(partial, nb, partial2);
      compress(partial + nb, nb, partial2 + m);
      k_finalize2<<<1, kBlock, 0, stream>>>(partial2, m, scal, s0, s1);
    } else {
      k_finalize2<<<1, kBlock, 0, stream>>>(partial, nb, scal, s0, s1);
    }
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
  // z-march layout of a 3-D level of 2^18 cells and more: about 8192
  // blocks of kZmX x kZmY columns (the best of 1024..16384 blocks and 4 or
  // 8 rows measured at 256^3; never more than ceil(n / kBlock), the partial
  // buffer's share); 0 when the level does not qualify
  long zm_layout(const Level<double> &l, dim3 &grid, int &chunk) const {
    if (ndim != 3 || l.n < (1L << 18)) return 0;
    const long nb = ceil_div(l.n, kBlock);
    const long bxy = (long)ceil_div(l.nx, kZmX) * ceil_div(l.ny, kZmY);
    if (bxy > nb || ceil_div(l.ny, kZmY) > 65535) return 0;
    const long slabs = std::max(1L, std::min(8192L, nb) / bxy);
    chunk = (int)ceil_div(l.nz, slabs);
    grid = dim3(ceil_div(l.nx, kZmX), ceil_div(l.ny, kZmY), ceil_div(l.nz, chunk));
    return bxy * grid.z;
  }
  // b - A x on an FP64 level (res stored when asked), r.r partials in
  // partial; returns their number
  int residual_partials(const Level<double> &l, bool store, const double *flag = nullptr) {
    dim3 grid;
    int chunk = 1;
    const long nz = zm_layout(l, grid, chunk);
    if (nz > 0) {
      if (variable) k_residual_zm<true><<<grid, kZmX * kZmY, 0, stream>>>(view(l, false), store ? 1 : 0, partial,
                                                                          flag, chunk);
      else k_residual_zm<false><<<grid, kZmX * kZmY, 0, stream>>>(view(l, false), store ? 1 : 0, partial, flag,
                                                                  chunk);
      SCIHPC_CHECK_LAUNCH();
      return (int)nz;
    }
    const unsigned nb = ceil_div(l.n, kBlock);
    SCIHPC_DISPATCH(k_residual, double, nb, view(l, false), store ? 1 : 0, partial, flag);
    return (int)nb;
  }
  void reduce_residual(const Level<double> &l, bool store, const double *flag = nullptr) {
    finalize(residual_partials(l, store, flag), S_NORM);
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
      SCIHPC_DISPATCH(k_residual, double, nb, view(l, false), 1, partial, flag);
      finalize(nb, S_R2, 1);
      SCIHPC_DISPATCH(k_apply_dot, double, nb, view(l, false), l.res, l.buf, partial, flag);
      finalize(nb, S_RAR);
      k_sd_update<double><<<nb, kBlock, 0, stream>>>(view(l, false), scal, partial);
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
    k_child_avg<<<ceil_div(b.n, kBlock), kBlock, 0, stream>>>(view(a, false), view(b, false), b.deg / a.deg, ndim);
    SCIHPC_CHECK_LAUNCH();
    compat(level + 1, b.rhs);
  }

  void prolongation_ref(int level) {
    auto &c = d[level], &f = d[level - 1];
    k_slope_prolong<<<ceil_div(f.n, kBlock), kBlock, 0, stream>>>(view(c, false), view(f, false), c.deg / f.deg, ndim, c.dx,
                                                                  c.dy, c.dz, f.dx, f.dy, f.dz);
    SCIHPC_CHECK_LAUNCH();
    relax(level - 1, 10);
  }

  void coarse_solve_ref() {
    auto &c = d[L - 1];
    k_coarse_refine<<<1, kBlock, 0, stream>>>(view(c, false), Minv_ref, 0);
    SCIHPC_CHECK_LAUNCH();
    compat(L - 1, c.sol);
    double old_res = -1.0, new_res = 0.0;
    do {
      old_res = old_res < 0.0 ? residual_norm(L - 1, true) : new_res;
      k_coarse_refine<<<1, kBlock, 0, stream>>>(view(c, false), Minv_ref, 1);
      SCIHPC_CHECK_LAUNCH();
      compat(L - 1, c.sol);
      new_res = residual_norm(L - 1, true);
    } while (1.0 - new_res / old_res > 0.1);
  }

  // ------------------------------------------------ coarse inverse
  // dense inverse of the coarsest operator into `dst`; rb selects the
  // red-black solver's operator (Dirichlet coarse wall term) over the
  // reference's
  void build_coarse_inverse(bool rb, double *dst) {
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
            if (bc == SCIHPC_MG_DIRICHLET && rb && c.deg > 1) {
              const DLevel<double> v = view(c, true);
              const int walls_x = (i == 0) + (i == c.nx - 1), walls_y = (j == 0) + (j == c.ny - 1);
              const int walls_z = ndim == 3 ? (kk == 0) + (kk == c.nz - 1) : 0;
              diag = -(cc + walls_x * v.wx + walls_y * v.wy + walls_z * v.wz);
            } else {
              diag = bc == SCIHPC_MG_DIRICHLET
                         ? (ndim == 2 ? -2.0 * ax - 2.0 * ay : -2.0 * ax - 2.0 * ay - 2.0 * az)
                         : -cc;
            }
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
    SCIHPC_HIP(hipMemcpyAsync(dst, Inv.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, stream));
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
      b.sol_alt = dalloc<float>(b.n, stream);
      b.rhs = dalloc<float>(b.n, stream);
      b.rhs_alt = dalloc<float>(b.n, stream);
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
    if (l == 0 && rho0) {
      if (!rho0f) rho0f = dalloc<float>(a.n, stream);
      k_to_float<<<nb, kBlock, 0, stream>>>(rho0, rho0f, a.n);
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
    build_coarse_inverse(true, Minv);  // Neumann: Minv_ref aliases Minv
  }

  // ------------------------------------------------------- solve
  // FP64 work vectors of the flexible-CG solve
  double *fx_ = nullptr, *fr_ = nullptr, *fp_[2] = {nullptr, nullptr}, *fq_ = nullptr;
  double *fr2_ = nullptr;  // second residual buffer of the fused update (lazy)
  double last_rr_ = 0.0;  // final relative residual of the last flexible-CG solve
  // lagged convergence checks (opts.lagged_check): per path (stationary or
  // flexible CG, FP64 or mixed) the iteration count of the last converged
  // solve, with the tolerance and smoothing counts it ran with
  int lag_iters_[4] = {0, 0, 0, 0};
  double lag_tol_[4] = {0.0, 0.0, 0.0, 0.0};
  int lag_nu_[4] = {-1, -1, -1, -1};
  int first_check_ = 0;  // first iteration whose residual the last solve read
  void ensure_fcg() {
    if (fx_) return;
    const long n = d[0].n;
    fx_ = dalloc<double>(n, stream);
    fr_ = dalloc<double>(n, stream);
    fp_[0] = dalloc<double>(n, stream);
    fp_[1] = dalloc<double>(n, stream);
    fq_ = dalloc<double>(n, stream);
  }

  // Krylov-accelerated defect correction: FCG(1) with the V-cycle (FP64 or
  // FP32) as preconditioner; x, r, p, q and every inner product in FP64.
  // learned: a guess is being formed on lstream; after the first iteration
  // it replaces that iteration if its FP64 residual is lower
  int solve_fcg(const scihpc_solve_opts &o, scihpc_solve_stats &st, bool zero, double bnorm, double scale,
                bool learned_overlap, int check_from) {
    ensure_fcg();
    auto &l0 = d[0];
    const long n = l0.n;
    const unsigned nb = ceil_div(n, kBlock);
    const bool mixed = o.precision == SCIHPC_PREC_MIXED;
    if (mixed) ensure_fp32();
    double *x = fx_;
    // a singular all-Neumann operator solved below 1e-9: the step sizes
    // leave out the constant component of the direction (fcg_dir_sums),
    // which otherwise drives the iteration away near the FP64 floor (above
    // it the final check of b - A x below covers the drift)
    const double corr_n = bc == SCIHPC_MG_NEUMANN && o.rel_tol < 1e-9 ? double(n) : 0.0;
    const bool sums = corr_n > 0.0;
    // r = b - A x (x = 0, or the current level-0 solution); the FP64
    // preconditioner works in place on level 0, whose right-hand side *is*
    // r during the solve (b waits in fr_)
    bool swapped = false;
    auto start = [&](bool from_zero) -> double {
      if (swapped) {
        std::swap(l0.rhs, fr_);
        swapped = false;
      }
      if (from_zero) {
        SCIHPC_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, stream));
        k_copy_cast<<<nb, kBlock, 0, stream>>>(l0.rhs, fr_, mixed ? f[0].rhs : nullptr, n);
        SCIHPC_CHECK_LAUNCH();
      } else {
        SCIHPC_HIP(hipMemcpyAsync(x, l0.sol, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
        reduce_residual(l0, true);
        k_copy_cast<<<nb, kBlock, 0, stream>>>(l0.res, fr_, mixed ? f[0].rhs : nullptr, n);
        SCIHPC_CHECK_LAUNCH();
      }
      if (!mixed) {
        std::swap(l0.rhs, fr_);
        swapped = true;
      }
      return from_zero ? (bnorm > 0.0 ? 1.0 : 0.0) : std::sqrt(fetch(S_NORM)) / scale;
    };
    double rr = start(zero);
    double *r = mixed ? fr_ : l0.rhs;
    int it = 0, cur = 0, spent = 0;
    bool corr_left = mixed && o.learned_correction != 0 && learned.active();
    // 3-D levels of 2^18 cells and more: the direction kernel marches in z
    dim3 zm_grid(1, 1, 1);
    int zm_chunk = 1;
    const long nzm = zm_layout(l0, zm_grid, zm_chunk);
    const bool zm = nzm > 0;
    const unsigned nbf = ceil_div(n, kFcgThreads);  // workgroups of the FCG vector kernels
    const int ndir = zm ? (int)nzm : (int)nbf;
    // mixed 2-D solves on a tiled level 0: the update of x and r runs inside
    // the next V-cycle's level-0 pre-smoothing (k_pre2d_fcg), r alternating
    // between two buffers; never while a learned step is pending
    const bool fuse_upd = mixed && fused_update && !zm && o.nu1 > 0 && o.nu1 <= 2 && o.nu2 <= 2 &&
                          fused_level(f, 0) && block_level(f) > 0 &&
                          (long)tiles2d(f[0], 0) <= 2 * (long)nb;
    if (fuse_upd && !fr2_) fr2_ = dalloc<double>(n, stream);
    bool pre0_done = false;
    for (;;) {
    while (rr > o.rel_tol && it + spent < o.max_cycles) {
      // z = B r; from the second iteration on the level-0 post kernel also
      // forms z . q_old (S_ZQ) for the flexible direction update
      OuterArgs oa{};
      oa.q = fq_;
      const void *z;
      int np;
      if (mixed) {
        np = vcycle(f, o.nu1, o.nu2, true, it > 0 ? 3 : 0, oa, pre0_done);
        z = f[0].sol;
      } else {
        np = vcycle(d, o.nu1, o.nu2, true, it > 0 ? 3 : 0, oa);
        z = l0.sol;
      }
      // small reductions are folded into the consumers: z.q_old into the
      // direction kernel, p.q and p.r into the update (its own partials go
      // to [0, nbf), the direction's to [2 nb, 2 nb + 2 ndir))
      // (z.q_old partials at [0, np) must end before the direction's)
      const bool fuse = nbf <= kCompressAbove && np <= kCompressAbove && np <= 2 * (int)nb && ndir <= (int)nb;
      if (it > 0 && (!fuse || zm)) finalize(np, S_ZQ);
      double *pold = fp_[cur], *pnew = fp_[cur ^ 1];
      double *dpart = fuse ? partial + 2 * (long)nb : partial;
      const double *zq_part = (fuse && it > 0) ? partial : nullptr;
      const int zq_n = (fuse && it > 0) ? np : 0;
#define SCIHPC_FCG_DIR_S(T, D, V, S)                                                                     \
  k_fcg_direction<T, D, V, S><<<nbf, kFcgThreads, 0, stream>>>(view(l0), (const T *)z, pold, pnew, fq_, r, scal, \
                                                               it == 0, dpart, zq_part, zq_n)
#define SCIHPC_FCG_DIR(T, D, V)                   \
  do {                                            \
    if (sums) SCIHPC_FCG_DIR_S(T, D, V, true);    \
    else SCIHPC_FCG_DIR_S(T, D, V, false);        \
  } while (0)
#define SCIHPC_FCG_ZM_S(T, V, S)                                                                            \
  k_fcg_direction_zm<T, V, S><<<zm_grid, kZmX * kZmY, 0, stream>>>(view(l0), (const T *)z, pold, pnew, fq_, r, \
                                                                    scal, it == 0, dpart, zm_chunk)
#define SCIHPC_FCG_ZM(T, V)                    \
  do {                                         \
    if (sums) SCIHPC_FCG_ZM_S(T, V, true);     \
    else SCIHPC_FCG_ZM_S(T, V, false);         \
  } while (0)
      if (zm) {
        if (mixed) { if (variable) SCIHPC_FCG_ZM(float, true); else SCIHPC_FCG_ZM(float, false); }
        else { if (variable) SCIHPC_FCG_ZM(double, true); else SCIHPC_FCG_ZM(double, false); }
      } else if (mixed) {
        if (ndim == 2) { if (variable) SCIHPC_FCG_DIR(float, 2, true); else SCIHPC_FCG_DIR(float, 2, false); }
        else { if (variable) SCIHPC_FCG_DIR(float, 3, true); else SCIHPC_FCG_DIR(float, 3, false); }
      } else {
        if (ndim == 2) { if (variable) SCIHPC_FCG_DIR(double, 2, true); else SCIHPC_FCG_DIR(double, 2, false); }
        else { if (variable) SCIHPC_FCG_DIR(double, 3, true); else SCIHPC_FCG_DIR(double, 3, false); }
      }
#undef SCIHPC_FCG_ZM
#undef SCIHPC_FCG_ZM_S
#undef SCIHPC_FCG_DIR
#undef SCIHPC_FCG_DIR_S
      SCIHPC_CHECK_LAUNCH();
      if (!fuse) {
        k_fcg_dir_final<<<1, 1024, 0, stream>>>(dpart, ndir, scal, corr_n);
        SCIHPC_CHECK_LAUNCH();
      }
      cur ^= 1;
      const bool last = it + spent + 1 >= o.max_cycles;
      const bool fused_now = fuse_upd && fuse && !learned_overlap && !corr_left && !last;
      if (fused_now) {
        FcgUpdate u{x, r, r == fr_ ? fr2_ : fr_, fp_[cur], fq_, scal, dpart, ndir, corr_n, partial};
        pre2d_fcg(o.nu1, u);
        r = u.r_out;
      } else {
        k_fcg_update<<<nbf, kFcgThreads, 0, stream>>>(n, x, r, fp_[cur], fq_, scal, mixed ? f[0].rhs : nullptr,
                                                      partial, fuse ? dpart : nullptr, fuse ? ndir : 0, corr_n);
        SCIHPC_CHECK_LAUNCH();
      }
      pre0_done = fused_now;
      // the learned guess is enqueued while the first iteration runs
      if (learned_overlap && !learned_pending) learned_enqueue(bnorm);
      // lagged checks: iterations before check_from run without reading
      // the residual (the host keeps enqueueing); never while a learned
      // step is pending, and always on the last allowed iteration
      if (learned_overlap || corr_left || it + spent + 1 >= check_from || last) {
        if (!first_check_) first_check_ = it + spent + 1;
        rr = std::sqrt(host_sum(fused_now ? (int)tiles2d(f[0], 0) : (int)nbf)) / scale;
      }
      ++it;
      // the guess is judged at the end of the first iteration after which
      // its stream has finished (no wait): it replaces the iterate only if
      // its FP64 residual is below both ||b|| and the iterate's
      bool guess_ready = false;
      if (learned_overlap) {
        const hipError_t q = hipEventQuery(lev);
        if (q == hipErrorNotReady) (void)hipGetLastError();  // not an error: still running
        else SCIHPC_HIP(q);
        guess_ready = q == hipSuccess;
      }
      if (guess_ready) {
        learned_overlap = false;
        vec_norm2_cached_ = bnorm * bnorm;
        if (learned_combine(st, x, r)) {
          // restart from the combination: x = a x + c g, r = b - A x (mixed: b is l0.rhs)
          spent += it;
          it = 0;
          cur = 0;
          r = fr_;
          pre0_done = false;
          SCIHPC_HIP(hipMemcpyAsync(x, l0.sol, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
          reduce_residual(l0, true);
          k_copy_cast<<<nb, kBlock, 0, stream>>>(l0.res, fr_, f[0].rhs, n);
          SCIHPC_CHECK_LAUNCH();
          rr = std::sqrt(fetch(S_NORM)) / scale;
        }
      }
      // the learned coarse-level correction, once, after the first
      // iteration whose guess pipeline (if any) has been judged
      if (corr_left && !learned_pending && rr > o.rel_tol && it + spent < o.max_cycles) {
        corr_left = false;
        if (learned_correct(o.learned_correction, st, rr * scale, x)) {
          spent += it;
          it = 0;
          cur = 0;
          r = fr_;
          pre0_done = false;
          rr = std::sqrt(fetch(S_NORM)) / scale;
        }
      }
    }
    if (learned_overlap && learned_pending) {  // converged before the guess was ready: unused
      SCIHPC_HIP(hipEventSynchronize(lev));
      learned_pending = false;
      learned_overlap = false;
      st.learned_rejected++;
    }
    // r is a recurrence and drifts from b - A x by rounding: a solve it
    // calls converged is checked against the FP64 residual of x, and
    // continues from x (r = b - A x afresh) while that is above rel_tol and
    // iterations remain; the reported residual is always b - A x's (one
    // residual pass per solve)
    if (swapped) {
      std::swap(l0.rhs, fr_);
      swapped = false;
    }
    const bool recurrence_converged = rr <= o.rel_tol;
    SCIHPC_HIP(hipMemcpyAsync(l0.sol, x, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
    reduce_residual(l0, false);
    rr = std::sqrt(fetch(S_NORM)) / scale;
    if (!recurrence_converged || rr <= o.rel_tol || it + spent >= o.max_cycles) break;
    spent += it;
    it = 0;
    cur = 0;
    pre0_done = false;
    corr_left = false;
    rr = start(false);
    r = mixed ? fr_ : l0.rhs;
    }
    it += spent;
    last_rr_ = rr;
    if (swapped) std::swap(l0.rhs, fr_);  // restore b
    SCIHPC_HIP(hipMemcpyAsync(l0.sol, x, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
    if (bc == SCIHPC_MG_NEUMANN) sub_mean(l0.sol, n);
    st.cycles = it;
    return it;
  }

  int solve(const scihpc_solve_opts &o, scihpc_solve_stats &st) {
    SCIHPC_REQUIRE(o.rel_tol > 0 && o.max_cycles >= 0 && o.nu1 >= 0 && o.nu2 >= 0, "bad solve options");
    SCIHPC_REQUIRE(o.nu1 <= 2 && o.nu2 <= 2, "nu1, nu2 <= 2 (tiled kernels stage a 5-cell halo)");
    SCIHPC_REQUIRE(o.accel == 0 || o.accel == 1, "accel must be 0 (stationary) or 1 (flexible CG)");
    SCIHPC_REQUIRE(o.lagged_check == 0 || o.lagged_check == 1, "lagged_check must be 0 or 1");
    SCIHPC_REQUIRE(o.use_learned >= 0 && o.use_learned <= 3, "use_learned must be 0..3");
    SCIHPC_REQUIRE(o.learned_correction >= 0 && o.learned_correction <= 2, "learned_correction must be 0..2");
    SCIHPC_REQUIRE(o.learned_correction == 0 || (o.accel == 1 && o.precision == SCIHPC_PREC_MIXED),
                   "learned_correction needs a mixed-precision flexible-CG solve (accel = 1)");
    auto &l0 = d[0];
    const unsigned nb = ceil_div(l0.n, kBlock);
    SCIHPC_HIP(hipEventRecord(ev0, stream));
    const double bnorm = std::sqrt(vec_norm2(l0.rhs, l0.n));
    const double scale = bnorm > 0.0 ? bnorm : 1.0;
    st = scihpc_solve_stats{};
    int cycles = 0;
    double rr;
    bool zero = o.zero_guess != 0;
    bool overlap = false;
    bool gated = false;
    // lagged checks: the residual is first read one iteration before the
    // count that the last converged solve of the same path needed
    const int key = (o.accel == 1 ? 2 : 0) + (o.precision == SCIHPC_PREC_MIXED ? 1 : 0);
    const int nu = o.nu1 * 16 + o.nu2;
    const int check_from =
        o.lagged_check && lag_tol_[key] == o.rel_tol && lag_nu_[key] == nu ? lag_iters_[key] - 1 : 0;
    first_check_ = 0;
    if (o.use_learned == 1 && learned.active() && learned_skip > 0) {
      --learned_skip;
      st.learned_skipped = 1;
    } else if (o.use_learned && learned.active() && bnorm > 0.0) {
      // use_learned == 2 keeps the guess unjudged (ablation measurements);
      // 1 and 3 judge it (1 with the back-off)
      gated = o.use_learned == 1;
      overlap = o.use_learned != 2 && o.accel == 1 && o.precision == SCIHPC_PREC_MIXED && o.max_cycles > 0;
      if (!overlap) {
        learned_enqueue(bnorm);
        if (learned_decide(st, bnorm, 1.0, o.use_learned == 2)) {
          zero = false;
        } else {  // fall back to plain multigrid from p = 0
          SCIHPC_HIP(hipMemsetAsync(l0.sol, 0, sizeof(double) * l0.n, stream));
          zero = true;
        }
      }
    }
    if (o.accel == 1) {
      solve_fcg(o, st, zero, bnorm, scale, overlap, check_from);
      cycles = st.cycles;
      rr = last_rr_;
    } else if (o.precision == SCIHPC_PREC_FP64) {
      rr = zero ? (bnorm > 0.0 ? 1.0 : 0.0) : residual_norm(0, false) / scale;
      while (rr > o.rel_tol && cycles < o.max_cycles) {
        const int np = vcycle(d, o.nu1, o.nu2, zero && cycles == 0, 1);
        if (cycles + 1 >= check_from || cycles + 1 >= o.max_cycles) {
          if (!first_check_) first_check_ = cycles + 1;
          rr = std::sqrt(host_sum(np)) / scale;
        }
        ++cycles;
      }
      if (zero && cycles == 0) SCIHPC_HIP(hipMemsetAsync(l0.sol, 0, sizeof(double) * l0.n, stream));
      if (bc == SCIHPC_MG_NEUMANN) sub_mean(l0.sol, l0.n);
    } else {
      ensure_fp32();
      auto &f0 = f[0];
      // p: l0.sol / p_alt (ping-pong); FP32 rhs0 = float(b - A p)
      double *p = l0.sol, *q = p_alt;
      auto first = [&](int is_first) {
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
        std::swap(p, q);
        return std::sqrt(host_sum(nb)) / scale;
      };
      // initial residual into the FP32 right-hand side (e = 0)
      if (!zero) SCIHPC_HIP(hipMemsetAsync(f0.sol, 0, sizeof(float) * f0.n, stream));
      rr = first(zero ? 1 : 0);
      while (rr > o.rel_tol && cycles < o.max_cycles) {
        OuterArgs oa{};
        oa.P = view(l0);
        oa.p_old = p;
        oa.p_new = q;
        oa.r32 = f0.rhs_alt;
        const int np = vcycle(f, o.nu1, o.nu2, true, 2, oa);
        std::swap(f0.rhs, f0.rhs_alt);
        std::swap(p, q);
        if (cycles + 1 >= check_from || cycles + 1 >= o.max_cycles) {
          if (!first_check_) first_check_ = cycles + 1;
          rr = std::sqrt(host_sum(np)) / scale;
        }
        ++cycles;
      }
      if (p != l0.sol) {
        // keep the result in l0.sol; p_alt becomes the spare buffer
        std::swap(l0.sol, p_alt);
      }
      if (bc == SCIHPC_MG_NEUMANN) sub_mean(l0.sol, l0.n);
    }
    SCIHPC_HIP(hipEventRecord(ev1, stream));
    SCIHPC_HIP(hipEventSynchronize(ev1));
    float ms = 0.f;
    SCIHPC_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    st.cycles = cycles;
    st.rel_residual = rr;
    st.ms = ms;
    if (o.lagged_check && rr <= o.rel_tol && cycles > 0) {
      // converged at the first residual read: the count may overshoot, so
      // the next solve reads one iteration earlier
      lag_iters_[key] = first_check_ == cycles && check_from > 0 ? cycles - 1 : cycles;
      lag_tol_[key] = o.rel_tol;
      lag_nu_[key] = nu;
    }
    if (gated) learned_note(st);
    return rr <= o.rel_tol ? SCIHPC_OK : SCIHPC_NOT_CONVERGED;
  }

  // every cell of level l is the block of deg^ndim level-0 cells
  bool block_coarsening(int l) const {
    const auto &a = d[l];
    return a.nx * a.deg == d[0].nx && a.ny * a.deg == d[0].ny && (ndim == 2 || a.nz * a.deg == d[0].nz);
  }
  // learned features of the current b (and sol when tgt) on the hook's level
  void learned_features(int lvl, double sigma, float *feat, float *tgt) {
    auto &a = d[lvl];
    const int bs = a.deg;
    k_learned_features<<<ceil_div(a.n, kBlock), kBlock, 0, stream>>>(
        d[0].rhs, variable ? rho0 : nullptr, tgt ? d[0].sol : nullptr, d[0].nx, d[0].ny, bs, ndim == 3 ? bs : 1, a.nx,
        a.ny, ndim == 3 ? a.nz : 1, sigma > 0.0 ? 1.0 / sigma : 0.0, feat, tgt);
    SCIHPC_CHECK_LAUNCH();
  }

  // ------------------------------------------------ learned guess
  // p0 = alpha * S P S N(b, ln rho): N runs on the hook's level, S is two
  // red-black sweeps on a level (against the child-averaged b), P prolongs
  // one level up, alpha = (p.b)/(p.A p) is the energy-optimal scale.  The
  // pipeline is enqueued on its own stream with its own reduction buffers,
  // so in the mixed flexible-CG solve (whose V-cycle works on the FP32
  // levels) it runs while the iterations from p = 0 run; when it is ready the
  // guess replaces the iterate only if its FP64 residual is below both ||b||
  // and the iterate's.  Otherwise (FP64 V-cycles share the FP64 levels) it
  // runs first and is kept only if its residual is below ||b||.
  hipStream_t lstream = nullptr;
  hipEvent_t lev = nullptr;
  double *lpartial = nullptr, *lpartial2 = nullptr, *lscal = nullptr, *lhost = nullptr;
  bool learned_pending = false;
  double vec_norm2_cached_ = 0.0;  // ||b||^2 of the solve whose guess is pending
  // back-off of a guess that keeps losing (use_learned == 1): after the
  // k-th rejection in a row the next 2^k - 1 solves (at most 63) skip the
  // network; an accepted guess resets it
  int learned_streak = 0, learned_skip = 0;

  void learned_note(const scihpc_solve_stats &st) {
    if (st.learned_accepted > 0) {
      learned_streak = 0;
      learned_skip = 0;
    } else if (st.learned_rejected > 0) {
      learned_streak = std::min(learned_streak + 1, 6);
      learned_skip = (1 << learned_streak) - 1;
    }
  }

  void ensure_learned_buffers() {
    if (lstream) return;
    SCIHPC_HIP(hipStreamCreateWithFlags(&lstream, hipStreamNonBlocking));
    SCIHPC_HIP(hipEventCreateWithFlags(&lev, hipEventDisableTiming));
    const long nb = (long)ceil_div(d[0].n, kBlock);
    lpartial = dalloc<double>(4 * (nb + 64), stream);
    lpartial2 = dalloc<double>(4 * ((long)ceil_div(d[0].n, (long)kBlock * kBlock) + 64), stream);
    lscal = dalloc<double>(S_COUNT, stream);
    SCIHPC_HIP(hipHostMalloc(reinterpret_cast<void **>(&lhost), sizeof(double) * 4 * (nb + 64),
                             hipHostMallocDefault));
    SCIHPC_HIP(hipStreamSynchronize(stream));
  }
  void ensure_learned_io() {
    if (lfeat) return;
    const long hw = d[learned.level].n;
    lfeat = dalloc<float>(2 * hw, stream);
    lout = dalloc<float>(hw, stream);
    SCIHPC_HIP(hipStreamSynchronize(stream));
  }

  // enqueue the guess into l0.sol (and its four dots into lhost) on lstream
  // (b, rho and the FP64 faces are complete on the device when a solve
  // starts: the host wrote them synchronously and ||b|| was read back; the
  // main stream only reads them while the guess is formed)
  void learned_enqueue(double bnorm) {
    ensure_learned_buffers();
    ensure_learned_io();
    auto swap_all = [&] {
      std::swap(stream, lstream);
      std::swap(partial, lpartial);
      std::swap(partial2, lpartial2);
      std::swap(scal, lscal);
      std::swap(host_part, lhost);
    };
    swap_all();  // every helper below now enqueues on lstream with its buffers
    try {
      auto &l0 = d[0];
      const int lvl = learned.level;
      auto &a = d[lvl];
      const long hw = a.n;
      const double sigma = bnorm / std::sqrt((double)l0.n);
      learned_features(lvl, sigma, lfeat, nullptr);
      unet_infer(learned.net, lfeat, lout, stream);
      // b on the levels down to the network's (child averages of b, sol = 0)
      SCIHPC_HIP(hipMemsetAsync(l0.sol, 0, sizeof(double) * l0.n, stream));
      for (int l = 0; l < lvl; ++l) residual_restrict(d[l], d[l + 1]);
      k_net_to_level<<<ceil_div(hw, kBlock), kBlock, 0, stream>>>(lout, a.sol, hw,
                                                                   sigma * (double)unet_out_scale(learned.net));
      SCIHPC_CHECK_LAUNCH();
      sweep(a, 0, 0);
      sweep(a, 0, 0);
      for (int l = lvl - 1; l >= 0; --l) {
        SCIHPC_HIP(hipMemsetAsync(d[l].sol, 0, sizeof(double) * d[l].n, stream));
        prolong_add(d[l + 1], d[l]);
        sweep(d[l], 0, 0);
        sweep(d[l], 0, 0);
      }
      if (bc == SCIHPC_MG_NEUMANN) sub_mean(l0.sol, l0.n);
      const unsigned nb = ceil_div(l0.n, kBlock);
      if (ndim == 2) {
        if (variable) k_guess_dots<2, true><<<nb, kBlock, 0, stream>>>(view(l0), l0.sol, partial, (int)nb);
        else k_guess_dots<2, false><<<nb, kBlock, 0, stream>>>(view(l0), l0.sol, partial, (int)nb);
      } else {
        if (variable) k_guess_dots<3, true><<<nb, kBlock, 0, stream>>>(view(l0), l0.sol, partial, (int)nb);
        else k_guess_dots<3, false><<<nb, kBlock, 0, stream>>>(view(l0), l0.sol, partial, (int)nb);
      }
      SCIHPC_CHECK_LAUNCH();
      SCIHPC_HIP(hipMemcpyAsync(host_part, partial, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, stream));
    } catch (...) {
      swap_all();
      throw;
    }
    swap_all();
    SCIHPC_HIP(hipEventRecord(lev, lstream));
    learned_pending = true;
  }

  // The pending guess g (in d[0].sol) inside a mixed flexible-CG solve,
  // after k iterations from x = 0: the point of span{x_k, g} with the least
  // A(rho) energy error -- the error norm CG minimises -- solves
  // V^T A V c = V^T b for V = [x_k, g] (2 x 2; x_k.A x_k = x_k.b - x_k.r_k,
  // g.A x_k = g.b - g.r_k, g.b and g.A g from the guess pipeline).  It
  // replaces x_k (in d[0].sol, for the caller's restart) when its energy
  // J(v) = b.v - v.A v / 2 (A is negative definite: this is the SPD energy
  // of -A) is below x_k's by more than kCombineGain of |J(x_k)|.  J(x_k)
  // - J(x*) is half the squared energy error and |J(x_k)| is about half the
  // solution's squared energy norm, so the combination must remove at
  // least that fraction of the solution's energy from the error: enough to
  // be worth the restart it needs.  Late in a solve the iterate's error is
  // far below that and the guess is left; early, a guess better than the
  // iterate is taken.  A guess that cannot do that (NaN, zero, or no better
  // than the iterate) leaves x_k.
  // Returns true when replaced.
  bool learned_combine(scihpc_solve_stats &st, const double *x, const double *r) {
    double gb = 0.0, gag = 0.0, rel = -1.0;
    {
      SCIHPC_REQUIRE(learned_pending, "no learned guess pending");
      learned_pending = false;
      SCIHPC_HIP(hipEventSynchronize(lev));
      const unsigned nb = ceil_div(d[0].n, kBlock);
      double sg[4] = {0, 0, 0, 0};
      for (int k = 0; k < 4; ++k)
        for (unsigned q = 0; q < nb; ++q) sg[k] += lhost[k * nb + q];
      gb = sg[0];
      gag = sg[1];
      const double b2 = vec_norm2_cached_;
      const double alpha = gag != 0.0 ? gb / gag : 0.0;
      const double r2 = b2 - 2.0 * alpha * sg[2] + alpha * alpha * sg[3];
      rel = std::isfinite(r2) && b2 > 0.0 ? std::sqrt(std::fabs(r2) / b2) : -1.0;
    }
    st.learned_residual = rel;
    auto &l0 = d[0];
    const long n = l0.n;
    const unsigned nb = ceil_div(n, kBlock);
    SCIHPC_HIP(hipStreamWaitEvent(stream, lev, 0));
    k_combine_dots<<<nb, kBlock, 0, stream>>>(x, r, l0.rhs, l0.sol, n, partial, (int)nb);
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipMemcpyAsync(host_part, partial, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, stream));
    SCIHPC_HIP(hipStreamSynchronize(stream));
    double xb = 0.0, xr = 0.0, gr = 0.0;
    for (unsigned q = 0; q < nb; ++q) {
      xb += host_part[q];
      xr += host_part[nb + q];
      gr += host_part[2 * nb + q];
    }
    const double xax = xb - xr, xag = gb - gr;
    const double det = xax * gag - xag * xag;
    double a = 1.0, c = 0.0;
    bool ok = std::isfinite(det) && std::isfinite(gb) && gag != 0.0 && std::fabs(det) > 1e-14 * std::fabs(xax * gag);
    if (ok) {
      a = (xb * gag - xag * gb) / det;
      c = (xax * gb - xag * xb) / det;
      ok = std::isfinite(a) && std::isfinite(c);
    }
    auto energy = [&](double ca, double cc) {
      return ca * xb + cc * gb - 0.5 * (ca * ca * xax + 2.0 * ca * cc * xag + cc * cc * gag);
    };
    const double jk = energy(1.0, 0.0);
    const bool keep = ok && energy(a, c) < jk - kCombineGain * std::fabs(jk);
    if (!keep) {
      st.learned_rejected++;
      return false;
    }
    st.learned_accepted++;
    k_axpby<<<nb, kBlock, 0, stream>>>(x, l0.sol, n, a, c);
    SCIHPC_CHECK_LAUNCH();
    return true;
  }

  // judge the pending guess: kept if its relative FP64 residual is below
  // `threshold` (or always when `force`); a kept guess is scaled by alpha in
  // l0.sol on the main stream
  bool learned_decide(scihpc_solve_stats &st, double bnorm, double threshold, bool force) {
    double rel = -1.0;
    const bool keep = learned_judge(bnorm, threshold, force, rel);
    st.learned_residual = rel;
    if (keep && !force) st.learned_accepted++;
    if (!keep) st.learned_rejected++;
    return keep;
  }
  // the pending pipeline's output p against right-hand side g (||g|| =
  // gnorm, g in d[0].rhs when it ran): alpha = (p.g)/(p.A p) and rel =
  // ||g - alpha A p|| / ||g|| from its four inner products; when kept (rel <
  // threshold, or force) alpha p is left in d[0].sol on the main stream
  bool learned_judge(double gnorm, double threshold, bool force, double &rel) {
    SCIHPC_REQUIRE(learned_pending, "no learned guess pending");
    learned_pending = false;
    SCIHPC_HIP(hipEventSynchronize(lev));
    const unsigned nb = ceil_div(d[0].n, kBlock);
    double s[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (unsigned q = 0; q < nb; ++q) s[k] += lhost[k * nb + q];
    const double b2 = gnorm * gnorm;
    const double alpha = s[1] != 0.0 ? s[0] / s[1] : 0.0;
    const double r2 = b2 - 2.0 * alpha * s[2] + alpha * alpha * s[3];
    const bool finite = std::isfinite(alpha) && std::isfinite(r2) && alpha != 0.0;
    rel = finite && b2 > 0.0 ? std::sqrt(std::fabs(r2) / b2) : -1.0;
    if (finite && (force || rel < threshold)) {
      SCIHPC_HIP(hipStreamWaitEvent(stream, lev, 0));
      k_scale<<<nb, kBlock, 0, stream>>>(d[0].sol, d[0].n, alpha);
      SCIHPC_CHECK_LAUNCH();
      return true;
    }
    return false;
  }

  // The network as a coarse-level correction inside the mixed flexible-CG
  // refinement loop: the guess pipeline run on the current FP64 defect r
  // (r stands in for b in d[0].rhs while it runs), c = alpha S P S N(r),
  // judged by ||r - alpha A c|| < ||r||.  A kept correction moves x to
  // x + c and restarts the recurrence from r = b - A x (the directions so
  // far are conjugate for the old iterate only).  Returns true when kept.
  bool learned_correct(int mode, scihpc_solve_stats &st, double rnorm, double *x) {
    auto &l0 = d[0];
    double *b = l0.rhs;
    l0.rhs = fr_;
    double rel = -1.0;
    bool keep = false;
    try {
      learned_enqueue(rnorm);
      keep = learned_judge(rnorm, 1.0, mode == 2, rel);
    } catch (...) {
      l0.rhs = b;
      throw;
    }
    l0.rhs = b;
    st.correction_residual = rel;
    if (!keep) {
      st.correction_rejected++;
      return false;
    }
    if (mode != 2) st.correction_accepted++;
    const long n = l0.n;
    const unsigned nb = ceil_div(n, kBlock);
    k_add<<<nb, kBlock, 0, stream>>>(x, l0.sol, n);
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipMemcpyAsync(l0.sol, x, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
    reduce_residual(l0, true);
    k_copy_cast<<<nb, kBlock, 0, stream>>>(l0.res, fr_, f[0].rhs, n);
    SCIHPC_CHECK_LAUNCH();
    return true;
  }
};


// ================================================================ C ABI
namespace scihpc {
static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }
}  // namespace scihpc

extern "C" const char *scihpc_last_error(void) { return scihpc::g_last_error.c_str(); }

namespace {
std::mutex g_host_mu;
std::vector<void *> g_host_plain;  // blocks that fell back to calloc
}  // namespace

extern "C" int scihpc_host_alloc(size_t bytes, void **out) {
  return scihpc::guarded([&] {
    SCIHPC_REQUIRE(out, "out is null");
    *out = nullptr;
    const size_t n = bytes ? bytes : 1;
    void *p = nullptr;
    if (hipHostMalloc(&p, n, hipHostMallocDefault) == hipSuccess && p) {
      std::memset(p, 0, n);
    } else {
      (void)hipGetLastError();  // the refusal is not this call's error
      p = std::calloc(n, 1);
      SCIHPC_REQUIRE(p, "host allocation failed");
      std::lock_guard<std::mutex> lk(g_host_mu);
      g_host_plain.push_back(p);
    }
    *out = p;
  });
}

extern "C" int scihpc_host_free(void *p) {
  return scihpc::guarded([&] {
    if (!p) return;
    {
      std::lock_guard<std::mutex> lk(g_host_mu);
      auto it = std::find(g_host_plain.begin(), g_host_plain.end(), p);
      if (it != g_host_plain.end()) {
        g_host_plain.erase(it);
        std::free(p);
        return;
      }
    }
    SCIHPC_HIP(hipHostFree(p));
  });
}

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
      if (const char *e = std::getenv("SCIHPC_MG_BLOCK_MAX")) h->block_max = std::atol(e);
      if (const char *e = std::getenv("SCIHPC_MG_TILED")) h->tiled = std::atoi(e) != 0;
      if (const char *e = std::getenv("SCIHPC_MG_TILED3D")) h->tiled3d = std::atoi(e) != 0;
      if (const char *e = std::getenv("SCIHPC_MG_FUSED_UPDATE")) h->fused_update = std::atoi(e) != 0;
      if (const char *e = std::getenv("SCIHPC_MG_Z3_CHUNK")) h->z3_chunk = std::atoi(e);
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
        v.sol_alt = dalloc<double>(v.n, h->stream);
        v.rhs = dalloc<double>(v.n, h->stream);
        v.res = dalloc<double>(v.n, h->stream);
        v.buf = dalloc<double>(v.n, h->stream);
      }
      h->Minv = dalloc<double>((long)std::min<long>(h->d.back().n, kMaxCoarse) *
                               std::min<long>(h->d.back().n, kMaxCoarse), h->stream);
      h->partial = dalloc<double>(8 * ((long)ceil_div(h->d[0].n, kBlock) + 64), h->stream);
      h->partial2 = dalloc<double>(4 * ((long)ceil_div(h->d[0].n, (long)kBlock * kBlock) + 64), h->stream);
      SCIHPC_HIP(hipHostMalloc((void **)&h->host_part,
                               sizeof(double) * 4 * ((long)ceil_div(h->d[0].n, kBlock) + 64),
                               hipHostMallocDefault));
      h->scal = dalloc<double>(S_COUNT, h->stream);
      h->p_alt = dalloc<double>(h->d[0].n, h->stream);
      SCIHPC_HIP(hipHostMalloc((void **)&h->host_scal, sizeof(double) * S_COUNT, hipHostMallocDefault));
      {
        int dev = 0, lds = 0;
        SCIHPC_HIP(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
          h->lds_max = lds;
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0)
          h->num_cu = cu;
      }
      h->build_coarse_inverse(true, h->Minv);
      if (bc == SCIHPC_MG_DIRICHLET) {
        const long nc = std::min<long>(h->d.back().n, kMaxCoarse);
        h->Minv_ref = dalloc<double>(nc * nc, h->stream);
        h->build_coarse_inverse(false, h->Minv_ref);
      } else {
        h->Minv_ref = h->Minv;
      }
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
  if (h->lstream) (void)hipStreamSynchronize(h->lstream);
  for (hipEvent_t e : h->probe_ev) (void)hipEventDestroy(e);
  if (h->probe_clk) (void)hipFree(h->probe_clk);
  auto fr = [](void *p) {
    if (p) (void)hipFree(p);
  };
  for (auto &v : h->d) {
    fr(v.sol);
    fr(v.sol_alt);
    fr(v.rhs);
    fr(v.res);
    fr(v.buf);
    fr(v.cx);
    fr(v.cy);
    fr(v.cz);
  }
  for (auto &v : h->f) {
    fr(v.sol);
    fr(v.sol_alt);
    fr(v.rhs_alt);
    fr(v.rhs);
    fr(v.res);
    fr(v.cx);
    fr(v.cy);
    fr(v.cz);
  }
  if (h->Minv_ref != h->Minv) fr(h->Minv_ref);
  fr(h->Minv);
  fr(h->partial);
  fr(h->partial2);
  fr(h->blk_dev[0]);
  fr(h->blk_dev[1]);
  if (h->host_part) (void)hipHostFree(h->host_part);
  fr(h->fx_);
  fr(h->fr_);
  fr(h->fp_[0]);
  fr(h->fp_[1]);
  fr(h->fq_);
  fr(h->fr2_);
  fr(h->scal);
  fr(h->p_alt);
  fr(h->rho0);
  fr(h->rho0f);
  fr(h->lpartial);
  fr(h->lpartial2);
  fr(h->lscal);
  if (h->lhost) (void)hipHostFree(h->lhost);
  if (h->lev) (void)hipEventDestroy(h->lev);
  if (h->lstream) (void)hipStreamDestroy(h->lstream);
  fr(h->lfeat);
  fr(h->lout);
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
    if (!h->rho0) h->rho0 = dalloc<double>(n, h->stream);
    SCIHPC_HIP(hipMemcpyAsync(h->rho0, rho, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    h->set_faces_from_rho(h->rho0);
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
    if (!h->rho0) h->rho0 = dalloc<double>(l0.n, h->stream);
    try {
      SCIHPC_HIP(hipMemcpyAsync(dp, rho_padded, sizeof(double) * np, hipMemcpyHostToDevice, h->stream));
      k_padded_to_mg<<<ceil_div(l0.n, kBlock), kBlock, 0, h->stream>>>(dp, h->rho0, l0.nx, l0.ny, l0.nz, h->ndim);
      SCIHPC_CHECK_LAUNCH();
      h->set_faces_from_rho(h->rho0);
    } catch (...) {
      (void)hipFree(dp);
      throw;
    }
    SCIHPC_HIP(hipFree(dp));
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

// ------------------------------------------- residual without allocation
// scihpc_mg_residual_noalloc: the level-0 residual of scihpc_mg_residual on
// caller device arrays, with no handle and no allocation.  Its only device
// memory beyond the inputs and outputs is `scratch`: the per-block partial
// sums of ||r||^2 (one per kBlock cells) and their first-stage reduction.
// Every call on a GPU shares that GPU's copy; each call writes every entry
// it reads before reading it, so nothing has to start at zero.
namespace scihpc {
constexpr long kNoAllocPartials = 65536;                 // blocks: n <= 2^24 cells
constexpr long kNoAllocCells = kNoAllocPartials * kBlock;
constexpr int kNoAllocDevices = 64;
}  // namespace scihpc

static __device__ __align__(256) double scratch[scihpc::kNoAllocPartials + 512];

namespace scihpc {
// b - A x on level 0: A(rho) with the face coefficients formed on the fly
// (the expression and order of k_faces_from_rho, summed in stencil's order),
// or the constant operator of dv (rho == nullptr); res (nullable) written,
// partial[block] = sum of r^2 over the block's cells
template <int DIM, bool VAR>
__global__ void __launch_bounds__(kBlock)
    k_residual_free(DLevel<double> L, const double *__restrict__ rho, const double *__restrict__ x,
                    const double *__restrict__ b, double *__restrict__ res, double dx, double dy, double dz,
                    double *__restrict__ partial) {
  const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < L.n) {
    int i, j, k;
    cell_ijk(c, L.nx, L.ny, i, j, k);
    double off, diag;
    if constexpr (VAR) {
      const long sy = L.nx, sz = (long)L.nx * L.ny;
      const bool hl = i > 0, hr = i < L.nx - 1, hd = j > 0, hu = j < L.ny - 1;
      const double rc = rho[c];
      const double cR = hr ? 1.0 / (0.5 * (rc + rho[c + 1])) / dx / dx : 0.0;
      const double cL = hl ? 1.0 / (0.5 * (rho[c - 1] + rc)) / dx / dx : 0.0;
      const double cU = hu ? 1.0 / (0.5 * (rc + rho[c + sy])) / dy / dy : 0.0;
      const double cD = hd ? 1.0 / (0.5 * (rho[c - sy] + rc)) / dy / dy : 0.0;
      double o = 0.0;
      o += cL * x[hl ? c - 1 : c];
      o += cR * x[hr ? c + 1 : c];
      o += cD * x[hd ? c - sy : c];
      o += cU * x[hu ? c + sy : c];
      double d = -(cR + cL + cU + cD);
      if constexpr (DIM == 3) {
        const bool hb = k > 0, hf = k < L.nz - 1;
        const double cF = hf ? 1.0 / (0.5 * (rc + rho[c + sz])) / dz / dz : 0.0;
        const double cB = hb ? 1.0 / (0.5 * (rho[c - sz] + rc)) / dz / dz : 0.0;
        o += cB * x[hb ? c - sz : c];
        o += cF * x[hf ? c + sz : c];
        d -= (cF + cB);
      }
      off = o;
      diag = d;
    } else {
      stencil<double, DIM, false>(L, x, i, j, k, c, off, diag);
    }
    const double r = b[c] - (off + diag * x[c]);
    if (res) res[c] = r;
    v = r * r;
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// *out = sqrt(sum of partial[0..nb)), summed as k_finalize sums
__global__ void k_norm_out(const double *__restrict__ partial, int nb, double *__restrict__ out) {
  double s = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) s += partial[q];
  s = block_sum(s);
  if (threadIdx.x == 0) *out = sqrt(s);
}

// device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address)
static double *noalloc_scratch() {
  static std::mutex mu;
  static double *ptr[kNoAllocDevices] = {};
  int dev = 0;
  SCIHPC_HIP(hipGetDevice(&dev));
  SCIHPC_REQUIRE(dev >= 0 && dev < kNoAllocDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    SCIHPC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SCIHPC_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    ptr[dev] = static_cast<double *>(p);
  }
  return ptr[dev];
}
}  // namespace scihpc

extern "C" SCIHPC_API int scihpc_mg_residual_noalloc(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                                          int bc, const double *rho, const double *x, const double *b,
                                          double *res, double *norm, void *stream) {
  return guarded([&] {
    SCIHPC_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
    if (ndim == 2) {
      nz = 1;
      dz = 1.0;
    }
    SCIHPC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid sizes must be positive");
    SCIHPC_REQUIRE(dx > 0.0 && dy > 0.0 && dz > 0.0, "spacings must be positive");
    SCIHPC_REQUIRE(bc == SCIHPC_MG_NEUMANN || bc == SCIHPC_MG_DIRICHLET, "bc must be NEUMANN or DIRICHLET");
    SCIHPC_REQUIRE(rho == nullptr || bc == SCIHPC_MG_NEUMANN, "A(rho) needs SCIHPC_MG_NEUMANN walls");
    SCIHPC_REQUIRE(x && b && norm, "x, b and norm must be device pointers");
    const long n = (long)nx * ny * nz;
    if (n > kNoAllocCells)
      throw Error(SCIHPC_BAD_ARG, "scihpc_mg_residual_noalloc: " + std::to_string(n) + " cells exceed the " +
                                      std::to_string(kNoAllocCells) + " its fixed scratch array serves");
    double *part = noalloc_scratch();
    DLevel<double> L{};
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    L.n = n;
    L.ax = 1.0 / dx / dx;
    L.ay = 1.0 / dy / dy;
    L.az = ndim == 3 ? 1.0 / dz / dz : 0.0;
    L.dirichlet = bc == SCIHPC_MG_DIRICHLET;
    L.dconst = ndim == 2 ? -2.0 / dx / dx - 2.0 / dy / dy : -2.0 / dx / dx - 2.0 / dy / dy - 2.0 / dz / dz;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = ceil_div(n, kBlock);
    if (ndim == 2) {
      if (rho) k_residual_free<2, true><<<nb, kBlock, 0, s>>>(L, rho, x, b, res, dx, dy, dz, part);
      else k_residual_free<2, false><<<nb, kBlock, 0, s>>>(L, rho, x, b, res, dx, dy, dz, part);
    } else {
      if (rho) k_residual_free<3, true><<<nb, kBlock, 0, s>>>(L, rho, x, b, res, dx, dy, dz, part);
      else k_residual_free<3, false><<<nb, kBlock, 0, s>>>(L, rho, x, b, res, dx, dy, dz, part);
    }
    SCIHPC_CHECK_LAUNCH();
    const double *src = part;
    int cnt = (int)nb;
    if (nb > 2048) {  // first stage into the tail of scratch (at most 256 sums)
      const int m = (int)ceil_div(nb, kBlock);
      k_compress<<<m, kBlock, 0, s>>>(part, (int)nb, kBlock, part + kNoAllocPartials);
      SCIHPC_CHECK_LAUNCH();
      src = part + kNoAllocPartials;
      cnt = m;
    }
    k_norm_out<<<1, kBlock, 0, s>>>(src, cnt, norm);
    SCIHPC_CHECK_LAUNCH();
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
      if (variable) k_apply_dot<double, 2, true><<<nb, kBlock, 0, stream>>>(hh.view(l, false), l.res, l.buf, nullptr, nullptr);
      else k_apply_dot<double, 2, false><<<nb, kBlock, 0, stream>>>(hh.view(l, false), l.res, l.buf, nullptr, nullptr);
    } else {
      if (variable) k_apply_dot<double, 3, true><<<nb, kBlock, 0, stream>>>(hh.view(l, false), l.res, l.buf, nullptr, nullptr);
      else k_apply_dot<double, 3, false><<<nb, kBlock, 0, stream>>>(hh.view(l, false), l.res, l.buf, nullptr, nullptr);
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

extern "C" int scihpc_mg_attach_learned(scihpc_mg *h, scihpc_unet *net, int level) {
  return guarded([&] {
    check_level(h, 0);
    if (!net) {
      h->learned = scihpc::LearnedHook{};
      h->learned_streak = h->learned_skip = 0;
      return;
    }
    SCIHPC_REQUIRE(scihpc::unet_ndim(net) == h->ndim, "the network's dimension differs from the solver's");
    SCIHPC_REQUIRE(level >= 0 && level < h->L, "level out of range");
    const auto &a = h->d[level];
    SCIHPC_REQUIRE(scihpc::unet_width(net) == a.nx && scihpc::unet_height(net) == a.ny &&
                       (h->ndim == 2 || scihpc::unet_depth(net) == a.nz),
                   "network size differs from the level's grid");
    SCIHPC_REQUIRE(h->block_coarsening(level), "level is not a block coarsening");
    if (h->lfeat && (h->learned.level != level)) {
      SCIHPC_HIP(hipStreamSynchronize(h->stream));
      if (h->lstream) SCIHPC_HIP(hipStreamSynchronize(h->lstream));
      SCIHPC_HIP(hipFree(h->lfeat));
      SCIHPC_HIP(hipFree(h->lout));
      h->lfeat = h->lout = nullptr;
    }
    h->learned.net = net;
    h->learned.level = level;
    h->learned_streak = h->learned_skip = 0;
  });
}

extern "C" int scihpc_mg_learned_features(scihpc_mg *h, int level, float *feat, float *target, float *faces) {
  return guarded([&] {
    check_level(h, level);
    SCIHPC_REQUIRE(feat, "feat is null");
    SCIHPC_REQUIRE(!faces || h->variable, "faces need the variable operator (set_density)");
    const auto &a = h->d[level];
    SCIHPC_REQUIRE(h->block_coarsening(level), "level is not a block coarsening");
    const long hw = a.n;
    const int nd = h->ndim;
    const double bnorm = std::sqrt(h->vec_norm2(h->d[0].rhs, h->d[0].n));
    const double sigma = bnorm / std::sqrt((double)h->d[0].n);
    float *df = dalloc<float>(2 * hw, h->stream);
    float *dt = target ? dalloc<float>(hw, h->stream) : nullptr;
    float *dc = faces ? dalloc<float>((long)nd * hw, h->stream) : nullptr;
    try {
      h->learned_features(level, sigma, df, dt);
      SCIHPC_HIP(hipMemcpyAsync(feat, df, sizeof(float) * 2 * hw, hipMemcpyDeviceToHost, h->stream));
      if (target) SCIHPC_HIP(hipMemcpyAsync(target, dt, sizeof(float) * hw, hipMemcpyDeviceToHost, h->stream));
      if (faces) {
        k_learned_faces<<<ceil_div(hw, kBlock), kBlock, 0, h->stream>>>(
            a.cx, a.cy, a.cz, a.nx, a.ny, nd == 3 ? a.nz : 1, nd, a.dx * a.dx, a.dy * a.dy, a.dz * a.dz, dc);
        SCIHPC_CHECK_LAUNCH();
        SCIHPC_HIP(hipMemcpyAsync(faces, dc, sizeof(float) * nd * hw, hipMemcpyDeviceToHost, h->stream));
      }
      SCIHPC_HIP(hipStreamSynchronize(h->stream));
    } catch (...) {
      (void)hipFree(df);
      if (dt) (void)hipFree(dt);
      if (dc) (void)hipFree(dc);
      throw;
    }
    SCIHPC_HIP(hipFree(df));
    if (dt) SCIHPC_HIP(hipFree(dt));
    if (dc) SCIHPC_HIP(hipFree(dc));
  });
}

extern "C" int scihpc_mg_probe_kernel(scihpc_mg *h, int which, int capacity) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(which == -1 || which == 6, "which must be 6 (k_pre2d_fcg) or -1");
    SCIHPC_REQUIRE(capacity >= 0 && capacity <= 65536, "capacity must be 0..65536");
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    for (hipEvent_t e : h->probe_ev) SCIHPC_HIP(hipEventDestroy(e));
    h->probe_ev.clear();
    if (h->probe_clk) SCIHPC_HIP(hipFree(h->probe_clk));
    h->probe_clk = nullptr;
    h->probe_used = 0;
    h->probe_which = which;
    if (which < 0 || capacity == 0) return;
    // timing-only events: no system-scope fence on record, whose L2
    // writeback and invalidation would slow the bracketed kernel itself
    h->probe_ev.resize(2 * (size_t)capacity);
    for (auto &e : h->probe_ev) SCIHPC_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    // every tile kind of level 0 has at most one workgroup per 64 cells
    h->probe_stride = 2 * (long)ceil_div(h->d[0].n, 64);
    SCIHPC_HIP(hipMalloc(&h->probe_clk, sizeof(unsigned long long) * h->probe_stride * (size_t)capacity));
  });
}

extern "C" int scihpc_mg_probe_result(scihpc_mg *h, int *launches, double *total_ms, double *kernel_ms,
                                      double *bytes) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(h->probe_which >= 0 && h->probe_clk, "no probe armed");
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    double t = 0.0;
    for (int k = 0; k < h->probe_used; ++k) {
      float ms = 0.f;
      SCIHPC_HIP(hipEventElapsedTime(&ms, h->probe_ev[2 * k], h->probe_ev[2 * k + 1]));
      t += ms;
    }
    // the kernels' own spans: earliest workgroup start to latest workgroup
    // end on the wall clock (constant rate, kHz)
    const long ntiles = (long)h->tiles2d(h->f[0], 0);
    std::vector<unsigned long long> clk((size_t)h->probe_stride * h->probe_used);
    if (h->probe_used)
      SCIHPC_HIP(hipMemcpy(clk.data(), h->probe_clk, sizeof(unsigned long long) * clk.size(), hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    SCIHPC_HIP(hipGetDevice(&dev));
    SCIHPC_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    SCIHPC_REQUIRE(khz > 0, "the device reports no wall-clock rate");
    double span = 0.0;
    for (int k = 0; k < h->probe_used; ++k) {
      const unsigned long long *c = clk.data() + (size_t)h->probe_stride * k;
      unsigned long long t0 = ~0ull, t1 = 0;
      for (long b = 0; b < ntiles; ++b) {
        t0 = std::min(t0, c[2 * b]);
        t1 = std::max(t1, c[2 * b + 1]);
      }
      span += double(t1 - t0) / double(khz);
    }
    if (launches) *launches = h->probe_used;
    if (total_ms) *total_ms = t;
    if (kernel_ms) *kernel_ms = span;
    if (bytes) {  // the algorithmic bytes of one launch, as scihpc_mg_time_kernel(which = 6)
      const double n = (double)h->d[0].n, nc = h->L > 1 ? (double)h->d[1].n : 0.0;
      const double nf = h->variable ? h->ndim : 0;
      *bytes = (4.0 * nf + 56.0) * n + 8.0 * nc;
    }
    h->probe_used = 0;
  });
}

extern "C" int scihpc_mg_time_kernel(scihpc_mg *h, int which, int iters, double *ms, double *bytes) {
  return guarded([&] {
    check_level(h, 0);
    SCIHPC_REQUIRE(iters > 0 && which >= 0 && which <= 6, "bad timing request");
    SCIHPC_REQUIRE(h->L > 1, "timing needs at least two levels");
    h->ensure_fp32();
    auto &l0 = h->d[0];
    auto &f0 = h->f[0];
    const double n = (double)l0.n, nc = (double)h->d[1].n;
    const double nf = h->variable ? h->ndim : 0;  // face arrays per level
    const bool tiled = h->fused_level(h->d, 0);
    const bool z3 = !tiled && h->z3_level(h->d, 0) && h->d[1].deg == 2;
    // coefficient arrays a level-0 sweep reads: rho (z-streaming, A(rho)), faces, none
    const double na = z3 ? (h->variable ? 1.0 : 0.0) : nf;
    OuterArgs oa{};
    oa.P = h->view(l0);
    oa.p_old = l0.sol;
    oa.p_new = h->p_alt;
    oa.r32 = f0.rhs_alt;
    auto pre = [&](auto &lv) {
      for (int s = 0; s < 2; ++s) {
        if (z3) h->sweep3(lv, 0, 0, 0);
        else h->sweep(lv[0], 0, 0);
      }
      if (z3) h->rr3(lv, 0);
      else h->residual_restrict(lv[0], lv[1]);
    };
    auto post = [&](auto &lv, int mode) {
      h->prolong_add(lv[1], lv[0]);
      for (int s = 0; s < 2; ++s) {
        if (z3) h->sweep3(lv, 0, 1, 0);
        else h->sweep(lv[0], 1, 0);
      }
      h->post0_generic(lv, mode, oa);
    };
    // which = 6 runs on the vectors of the last mixed flexible-CG solve (x
    // moves by alpha p on every launch; the next solve starts afresh)
    FcgUpdate fu{};
    if (which == 6) {
      SCIHPC_REQUIRE(tiled && h->fr2_ != nullptr,
                     "which = 6 needs a tiled 2-D level 0 and a mixed flexible-CG solve first");
      const long nb = ceil_div(l0.n, kBlock);
      fu = FcgUpdate{h->fx_, h->fr_, h->fr2_, h->fp_[0], h->fq_, h->scal, h->partial + 2 * nb,
                     (int)ceil_div(l0.n, kFcgThreads), h->bc == SCIHPC_MG_NEUMANN ? double(l0.n) : 0.0,
                     h->partial};
    }
    auto launch = [&] {
      switch (which) {
        case 0:  // FP64 level-0 pre-smoothing (2 sweeps + residual + restriction)
          if (tiled) h->pre2d(h->d, 0, 2, 0);
          else pre(h->d);
          break;
        case 1:  // the same in FP32
          if (tiled) h->pre2d(h->f, 0, 2, 0);
          else pre(h->f);
          break;
        case 2:  // FP32 level-0 post-smoothing with the FP64 outer update
          if (tiled) h->post2d_mode<float, 2>(h->f, 0, 2, oa);
          else post(h->f, 2);
          break;
        case 3:  // FP64 level-0 post-smoothing with the residual norm
          if (tiled) h->post2d_mode<double, 1>(h->d, 0, 2, oa);
          else post(h->d, 1);
          break;
        case 4:  // FP64 prolongation of level 1 added to level 0, alone
          h->prolong_add(h->d[1], l0);
          break;
        case 5:  // FP64 level-0 residual partial sums (the norm's pass over the grid), alone
          h->residual_partials(l0, false);
          break;
        case 6:  // mixed FCG update fused into the FP32 level-0 pre-smoothing
          h->pre2d_fcg(2, fu);
          break;
      }
    };
    // algorithmic bytes: every array the launch touches, once per read and
    // once per write, at its stored precision
    double b = 0;
    if (which == 6) b = (4.0 * nf + 56.0) * n + 8.0 * nc;  // faces, FP64 r q x p, FP32 outputs
    else if (which == 4) b = 2.0 * 8.0 * n + 8.0 * nc;  // fine read + write, coarse read
    else if (which == 5) b = (2.0 + nf) * 8.0 * n;  // solution, right-hand side, faces
    else if (tiled) {  // one fused kernel per launch
      switch (which) {
        case 0: b = (3.0 + nf) * 8.0 * n + 2.0 * 8.0 * nc; break;
        case 1: b = (3.0 + nf) * 4.0 * n + 2.0 * 4.0 * nc; break;
        case 2: b = (2.0 + nf) * 4.0 * n + 4.0 * nc + (3.0 + nf) * 8.0 * n + 8.0 * n + 4.0 * n; break;
        case 3: b = (3.0 + nf) * 8.0 * n + 8.0 * nc; break;
      }
    } else {
      // sweep (3 + na) s N; residual + restriction (2 + na) s N + 2 s Nc;
      // prolong-add 2 s N + s Nc; FP64 outer update (32 + 8 nf) N;
      // FP64 residual norm (16 + 8 nf) N
      const double sd = which == 1 || which == 2 ? 4.0 : 8.0;
      if (which <= 1) b = 2.0 * (3.0 + na) * sd * n + (2.0 + na) * sd * n + 2.0 * sd * nc;
      else
        b = 2.0 * sd * n + sd * nc + 2.0 * (3.0 + na) * sd * n +
            (which == 2 ? (32.0 + 8.0 * nf) * n : (16.0 + 8.0 * nf) * n);
    }
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
