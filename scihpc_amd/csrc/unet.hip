// unet.hip -- the learned pressure guess: a 4-level U-Net (16/32/64/128
// channels, two 3x3 convolutions + ReLU per level, 2x2 average pooling,
// nearest upsampling with skip concatenation, 1x1 output head), its
// forward and training kernels (backward data, backward weight, loss, Adam),
// and the C ABI.  No SciHPC counterpart: the reference has no learned part.
//
// Storage: NHWC BF16 activations (pixel index = y*W + x, identical to the
// multigrid's x-fastest level layout), FP32 master parameters with BF16
// copies for the matrix cores, FP32 Adam moments.  Every layer of one
// parameter set is [w: cout x 9 x cin][b: cout] at its offset.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "common.hpp"
#include "learned.hpp"
#include "scihpc_hip.h"
#include "unet_kernels.hpp"

namespace scihpc {
namespace {

constexpr int kEB = 256;
constexpr int kLayers = 14;
constexpr int kHeadC = 16;  // channels entering the 1x1 head

// ------------------------------------------------------- elementwise
// Grids of a level: d x h x w (d = 1 in 2-D); a coarse cell (z, y, x) has
// the 2^DIM children (2z + a, 2y + b, 2x + c) on the finer level.
template <int DIM>
__device__ __forceinline__ long child0(long cp, int d, int h, int w) {  // first child of coarse pixel cp
  const int x = (int)(cp % w), y = (int)((cp / w) % h);
  const long r = cp / ((long)w * h);  // img * d + z
  const int z = (int)(r % d), img = (int)(r / d);
  const int D2 = DIM == 3 ? 2 * d : 1, z2 = DIM == 3 ? 2 * z : 0;
  return (((long)img * D2 + z2) * (2 * h) + 2 * y) * (2 * w) + 2 * x;
}
template <int DIM>
__device__ __forceinline__ long parent(long fp, int d, int h, int w) {  // coarse pixel of fine pixel fp
  const int W2 = 2 * w, H2 = 2 * h, D2 = DIM == 3 ? 2 * d : 1;
  const int x = (int)(fp % W2), y = (int)((fp / W2) % H2);
  const long r = fp / ((long)W2 * H2);
  const int z = (int)(r % D2), img = (int)(r / D2);
  return (((long)img * d + (DIM == 3 ? z / 2 : 0)) * h + y / 2) * w + x / 2;
}
template <int DIM>
__device__ __forceinline__ long child(long f0, int k, int h, int w) {  // k-th child (bits: x, y, z)
  return f0 + (k & 1) + ((k >> 1) & 1) * (2L * w) + (DIM == 3 ? ((k >> 2) & 1) * (4L * h * w) : 0L);
}

// The pooling and upsampling kernels move 8 channels (16 bytes) a thread:
// every channel count and view offset is a multiple of 16, so a thread's
// channels are one aligned uint4.  Per channel the arithmetic is the same
// as one element at a time.
constexpr int kVec = 8;
__device__ __forceinline__ void unpack8(uint4 v, float (&f)[kVec]) {
  const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    f[2 * q] = bf16_to_f((bf16_t)(u[q] & 0xffffu));
    f[2 * q + 1] = bf16_to_f((bf16_t)(u[q] >> 16));
  }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[kVec]) {
  unsigned u[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) u[q] = (unsigned)f_to_bf16(f[2 * q]) | ((unsigned)f_to_bf16(f[2 * q + 1]) << 16);
  return uint4{u[0], u[1], u[2], u[3]};
}
__device__ __forceinline__ uint4 &at8(View v, long p, int ch) {
  return *reinterpret_cast<uint4 *>(v.p + p * v.stride + v.off + ch);
}

// dst (coarse d x h x w) = mean of the children in src; thread per (pixel, 8 channels)
template <int DIM>
__global__ void __launch_bounds__(kEB) k_pool2(View src, View dst, int n, int d, int h, int w, int c) {
  const long e = (long)blockIdx.x * kEB + threadIdx.x;
  const int c8 = c / kVec;
  if (e >= (long)n * d * h * w * c8) return;
  const int ch = (int)(e % c8) * kVec;
  const long p = e / c8;
  const long f0 = child0<DIM>(p, d, h, w);
  float s[kVec] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < (1 << DIM); ++k) {
    float v[kVec];
    unpack8(at8(src, child<DIM>(f0, k, h, w), ch), v);
#pragma unroll
    for (int q = 0; q < kVec; ++q) s[q] += v[q];
  }
#pragma unroll
  for (int q = 0; q < kVec; ++q) s[q] = s[q] * (1.f / (1 << DIM));
  at8(dst, p, ch) = pack8(s);
}

// gf (fine) += gc[parent] / 2^DIM: gradient of k_pool2 added to the
// gradient that the same tensor already received through its skip
template <int DIM>
__global__ void __launch_bounds__(kEB) k_pool2_bwd_add(View gf, View gc, int n, int d, int h, int w, int c) {
  const long e = (long)blockIdx.x * kEB + threadIdx.x;
  const long nf = (long)n * (DIM == 3 ? 2 * d : 1) * (2 * h) * (2 * w);
  const int c8 = c / kVec;
  if (e >= nf * c8) return;
  const int ch = (int)(e % c8) * kVec;
  const long p = e / c8;
  const long pc = parent<DIM>(p, d, h, w);
  float a[kVec], b[kVec];
  unpack8(at8(gf, p, ch), a);
  unpack8(at8(gc, pc, ch), b);
#pragma unroll
  for (int q = 0; q < kVec; ++q) a[q] = a[q] + (1.f / (1 << DIM)) * b[q];
  at8(gf, p, ch) = pack8(a);
}

// dst (fine) = src[parent] (nearest upsampling)
template <int DIM>
__global__ void __launch_bounds__(kEB) k_up2(View src, View dst, int n, int d, int h, int w, int c) {
  const long e = (long)blockIdx.x * kEB + threadIdx.x;
  const long nf = (long)n * (DIM == 3 ? 2 * d : 1) * (2 * h) * (2 * w);
  const int c8 = c / kVec;
  if (e >= nf * c8) return;
  const int ch = (int)(e % c8) * kVec;
  const long p = e / c8;
  at8(dst, p, ch) = at8(src, parent<DIM>(p, d, h, w), ch);
}

// gc (coarse) = sum over the children of gf: gradient of k_up2
template <int DIM>
__global__ void __launch_bounds__(kEB) k_up2_bwd(View gf, View gc, int n, int d, int h, int w, int c) {
  const long e = (long)blockIdx.x * kEB + threadIdx.x;
  const int c8 = c / kVec;
  if (e >= (long)n * d * h * w * c8) return;
  const int ch = (int)(e % c8) * kVec;
  const long p = e / c8;
  const long f0 = child0<DIM>(p, d, h, w);
  float s[kVec] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < (1 << DIM); ++k) {
    float v[kVec];
    unpack8(at8(gf, child<DIM>(f0, k, h, w), ch), v);
#pragma unroll
    for (int q = 0; q < kVec; ++q) s[q] += v[q];
  }
  at8(gc, p, ch) = pack8(s);
}

// out[p] = b + sum_c x[p][c] w[c] (FP32 head weights)
__global__ void __launch_bounds__(kEB) k_head_fwd(View x, const float *__restrict__ wb, float *__restrict__ out,
                                                  long npix) {
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  if (p >= npix) return;
  float s = wb[kHeadC];
#pragma unroll
  for (int c = 0; c < kHeadC; ++c) s += bf16_to_f(x.p[p * x.stride + x.off + c]) * wb[c];
  out[p] = s;
}

// mean-squared error: dout = 2 (out - tgt) / npix, part[block] = sum (out - tgt)^2 / npix
__global__ void __launch_bounds__(kEB) k_mse(const float *__restrict__ out, const float *__restrict__ tgt,
                                             float *__restrict__ dout, long npix, double *__restrict__ part) {
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  double v = 0.0;
  if (p < npix) {
    const float d = out[p] - tgt[p];
    dout[p] = 2.0f * d / (float)npix;
    v = (double)d * (double)d / (double)npix;
  }
  v = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// energy loss of one image: E(e) / E(t) with e = out - tgt and
// E(v) = sum over faces c_f (v_i - v_j)^2 (the discrete A(rho) energy on the
// network's grid; face[p] = (c to x+1, c to y+1 (, c to z+1)), zero on
// walls); dout = 2 sum_f c_f (e_p - e_nb) / E(t) / n
template <int DIM>
__global__ void __launch_bounds__(kEB) k_energy_loss(const float *__restrict__ out, const float *__restrict__ tgt,
                                                     const float *__restrict__ face, const float *__restrict__ et,
                                                     float *__restrict__ dout, int n, int d, int h, int w,
                                                     double *__restrict__ part) {
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  double v = 0.0;
  if (p < (long)n * d * h * w) {
    const int x = (int)(p % w), y = (int)((p / w) % h);
    const long r = p / ((long)w * h);
    const int z = (int)(r % d), img = (int)(r / d);
    auto err = [&](long q) { return tgt ? out[q] - tgt[q] : out[q]; };
    const float e = err(p);
    const float inv = 1.f / (et[img] * (float)n);
    float g = 0.f;
    const int cnt[3] = {w, h, d};
    const int pos[3] = {x, y, z};
    const long step[3] = {1, (long)w, (long)w * h};
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      if (pos[a] + 1 < cnt[a]) {
        const float c = face[DIM * p + a];
        const float dd = e - err(p + step[a]);
        g += c * dd;
        v += (double)c * dd * dd * inv;
      }
      if (pos[a] > 0) g += face[DIM * (p - step[a]) + a] * (e - err(p - step[a]));
    }
    dout[p] = 2.f * g * inv;
  }
  v = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// One damped-Jacobi sweep of the network grid's A(rho) (the energy form
// E(v) = v.A v over the faces), applied to an error or, transposed, to a
// gradient:  M e = e - omega D^-1 A e,  M^T g = g - omega A (D^-1 g),
// (A v)_p = sum over p's faces c_f (v_p - v_nb), D_p = sum of p's faces.
// With smoothing sweeps S the energy loss is E(M^S e) / E(t): the error the
// multigrid smoother leaves of the network's output, so the network is not
// asked for high-frequency accuracy that the smoother provides anyway.
template <int DIM>
__device__ __forceinline__ float diag_of(const float *__restrict__ face, long p, const int (&pos)[3],
                                         const int (&cnt)[3], const long (&step)[3]) {
  float dsum = 0.f;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    if (pos[a] + 1 < cnt[a]) dsum += face[DIM * p + a];
    if (pos[a] > 0) dsum += face[DIM * (p - step[a]) + a];
  }
  return dsum;
}
template <int DIM>
__global__ void __launch_bounds__(kEB) k_jacobi(const float *__restrict__ v, float *__restrict__ o,
                                                const float *__restrict__ face, int n, int d, int h, int w,
                                                float omega, int transpose) {
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  if (p >= (long)n * d * h * w) return;
  const int pos[3] = {(int)(p % w), (int)((p / w) % h), (int)((p / ((long)w * h)) % d)};
  const int cnt[3] = {w, h, d};
  const long step[3] = {1, (long)w, (long)w * h};
  const float dp = diag_of<DIM>(face, p, pos, cnt, step);
  float acc = 0.f;
  if (!transpose) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      if (pos[a] + 1 < cnt[a]) acc += face[DIM * p + a] * (v[p] - v[p + step[a]]);
      if (pos[a] > 0) acc += face[DIM * (p - step[a]) + a] * (v[p] - v[p - step[a]]);
    }
    o[p] = dp > 0.f ? v[p] - omega * acc / dp : v[p];
  } else {
    const float up = dp > 0.f ? v[p] / dp : 0.f;
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int sgn = 0; sgn < 2; ++sgn) {
        const bool has = sgn ? pos[a] > 0 : pos[a] + 1 < cnt[a];
        if (!has) continue;
        const long q = sgn ? p - step[a] : p + step[a];
        int qpos[3] = {pos[0], pos[1], pos[2]};
        qpos[a] += sgn ? -1 : 1;
        const float dq = diag_of<DIM>(face, q, qpos, cnt, step);
        const float uq = dq > 0.f ? v[q] / dq : 0.f;
        acc += face[DIM * (sgn ? q : p) + a] * (up - uq);
      }
    o[p] = v[p] - omega * acc;
  }
}

// e = out - tgt
__global__ void __launch_bounds__(kEB) k_sub(const float *__restrict__ a, const float *__restrict__ b,
                                             float *__restrict__ o, long n) {
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  if (p < n) o[p] = a[p] - b[p];
}

// head backward: gx[p][c] = dout[p] w[c] (BF16); part[block][c] = sum_p dout x[p][c], part[block][16] = sum dout
__global__ void __launch_bounds__(kEB) k_head_bwd(const float *__restrict__ dout, View x, const float *__restrict__ wb,
                                                  View gx, float *__restrict__ part, long npix) {
  __shared__ float red[kEB / 64][kHeadC + 1];
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  const float g = p < npix ? dout[p] : 0.f;
  float v[kHeadC + 1];
#pragma unroll
  for (int c = 0; c < kHeadC; ++c) {
    float xv = 0.f;
    if (p < npix) {
      xv = bf16_to_f(x.p[p * x.stride + x.off + c]);
      gx.p[p * gx.stride + gx.off + c] = f_to_bf16(g * wb[c]);
    }
    v[c] = g * xv;
  }
  v[kHeadC] = g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c <= kHeadC; ++c) {
    const float s = wave_sum(v[c]);
    if (lane == 0) red[wave][c] = s;
  }
  __syncthreads();
  if (threadIdx.x <= kHeadC) {
    float s = 0.f;
    for (int wv = 0; wv < kEB / 64; ++wv) s += red[wv][threadIdx.x];
    part[(long)blockIdx.x * (kHeadC + 1) + threadIdx.x] = s;
  }
}

// dst[c] = sum_r part[r][c] in a fixed order (deterministic), coalesced: a
// block owns kSumCols adjacent columns, thread (x, y) sums rows y,
// y + kSumRowGroups, ... of column x, the kSumRowGroups partials are then
// added in row-group order
constexpr int kSumCols = 32, kSumRowGroups = 32;
__global__ void __launch_bounds__(kSumCols *kSumRowGroups) k_sum_rows_2d(const float *__restrict__ part, int rows,
                                                                         long cols, float *__restrict__ dst) {
  __shared__ float red[kSumRowGroups][kSumCols];
  const long c = (long)blockIdx.x * kSumCols + threadIdx.x;
  float s = 0.f;
  if (c < cols)
    for (int r = threadIdx.y; r < rows; r += kSumRowGroups) s += part[(long)r * cols + c];
  red[threadIdx.y][threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.y == 0 && c < cols) {
    float t = red[0][threadIdx.x];
    for (int y = 1; y < kSumRowGroups; ++y) t += red[y][threadIdx.x];
    dst[c] = t;
  }
}

// the same sums for a few columns: block c sums column c, each
// thread a fixed stride of rows, then a fixed-order tree (deterministic)
__global__ void __launch_bounds__(kEB) k_sum_rows_tree(const float *__restrict__ part, int rows, long cols,
                                                       float *__restrict__ dst) {
  __shared__ float red[kEB];
  const long c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += kEB) s += part[(long)r * cols + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kEB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) dst[c] = red[0];
}

__global__ void __launch_bounds__(kEB) k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                                              float *__restrict__ V, bf16_t *__restrict__ Pb, long n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * kEB + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  const float p = P[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  P[i] = p;
  Pb[i] = f_to_bf16(p);
}

__global__ void __launch_bounds__(kEB) k_to_bf16(const float *__restrict__ a, bf16_t *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * kEB + threadIdx.x;
  if (i < n) b[i] = f_to_bf16(a[i]);
}

// Wt[ci][T-1-t][co] = W[co][t][ci]: the backward-data convolution is the
// forward kernel on the gradient with these weights
__global__ void __launch_bounds__(kEB) k_flip(const bf16_t *__restrict__ W, bf16_t *__restrict__ Wt, int cin,
                                              int cout, int T) {
  const long e = (long)blockIdx.x * kEB + threadIdx.x;
  if (e >= (long)cout * T * cin) return;
  const int ci = (int)(e % cin);
  const int t = (int)((e / cin) % T);
  const int co = (int)(e / ((long)cin * T));
  Wt[((long)ci * T + (T - 1 - t)) * cout + co] = W[e];
}

// network input: channels (feat0, feat1, 0 ... 0) from FP32 features
// [sample][pixel][2]; optional sample indices and mirror bits per batch
// entry (1 = x, 2 = y, 4 = z); optional targets tgt =
// store_tgt[sample][pixel] * inv_scale and faces of the mirrored grid (a
// face between cells x and x+1 maps to W-2-x under an x mirror; the wall
// face stays zero) with the per-image energy norm
template <int DIM>
__global__ void __launch_bounds__(kEB) k_make_input(const float *__restrict__ feat, const float *__restrict__ tgt_store,
                                                    const float *__restrict__ face_store,
                                                    const float *__restrict__ e_store, const int *__restrict__ idx,
                                                    bf16_t *__restrict__ in, float *__restrict__ tgt,
                                                    float *__restrict__ face, float *__restrict__ et, int n, int d,
                                                    int h, int w, float inv_scale) {
  const long p = (long)blockIdx.x * kEB + threadIdx.x;
  if (p >= (long)n * d * h * w) return;
  const int x = (int)(p % w), y = (int)((p / w) % h);
  const long r = p / ((long)w * h);
  const int z = (int)(r % d), img = (int)(r / d);
  int s = img, fl = 0;
  if (idx) {
    s = idx[2 * img];
    fl = idx[2 * img + 1];
  }
  const int pos[3] = {x, y, z}, cnt[3] = {w, h, d};
  int src[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) src[a] = (fl >> a) & 1 ? cnt[a] - 1 - pos[a] : pos[a];
  const long base = (long)s * d * h * w;
  auto at = [&](int sx, int sy, int sz) { return base + ((long)sz * h + sy) * w + sx; };
  const long q = at(src[0], src[1], src[2]);
  bf16_t *o = in + p * 16;
  o[0] = f_to_bf16(feat[2 * q]);
  o[1] = f_to_bf16(feat[2 * q + 1]);
#pragma unroll
  for (int c = 2; c < 16; ++c) o[c] = 0;
  if (tgt) tgt[p] = tgt_store[q] * inv_scale;
  if (face) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      // face toward +1 along axis a of (x, y, z) in the mirrored image
      float c = 0.f;
      if (pos[a] + 1 < cnt[a]) {
        int sf[3] = {src[0], src[1], src[2]};
        if ((fl >> a) & 1) sf[a] = cnt[a] - 2 - pos[a];
        c = face_store[DIM * at(sf[0], sf[1], sf[2]) + a];
      }
      face[DIM * p + a] = c;
    }
    if (x == 0 && y == 0 && z == 0) et[img] = e_store[s];
  }
}

inline unsigned nblk(long n) { return ceil_div(n, kEB); }

// y = act(conv(x, w) + b) over s (2-D or 3-D); NB = s.cout / 16 output blocks
// Tiny images (fewer than 64 pixel tiles, a quarter of the CUs) split
// their output channels over workgroups, one or two 16-channel blocks
// each; larger ones keep every output block in one workgroup, which stages
// or reads each input tile once and reuses every A fragment for all its
// MFMAs.  With the LDS tile the split costs more than the parallelism it
// buys: at 256x512, batch 4, splitting every image under 1024 tiles gave a
// 0.485 ms forward pass, splitting under 64 gives 0.443 ms (the 32x64
// layers 46 -> 37 us, the 64x128 ones 25-90 -> 30-64 us).
constexpr long kSplitBelow = 64;
// the input tile goes through LDS (k_conv_lds) when it takes at most 96 KiB
// (2-D up to 224 input channels, 3-D up to 64) and at most six workgroups
// share it (an output-channel split over more would stage the same tile
// that many times: slower at 256x512's coarsest level, split 8); other
// shapes read it through the caches (k_conv).  Measured at 256x512, batch
// 4: the 64x128 layers 48 -> 25 us (32-64 input channels, split 4) and
// 101 -> 70 us (64 -> 192 channels, split 6), the full-resolution layers
// unchanged.
constexpr size_t kConvLdsMax = 98304;
constexpr int kConvLdsSplitMax = 6;
// weight-gradient workgroups aimed at per layer (the pixel chunk shrinks
// until there are about this many).  With the next pixels' loads in flight
// during the current ones' work, longer chunks pay: at 256x512, batch 4,
// 2048 / 1024 / 512 / 384 / 256 workgroups gave a training step of 1.57 /
// 1.52 / 1.47 / 1.51 / 1.63 ms (fewer partial rows to reduce, fewer
// epilogues, but below 512 too few workgroups to fill the CUs).
constexpr int kWgradWGs = 512;
template <int NB, int DIM, int G>
void conv_kernel_g(dim3 grid, bool lds, size_t bytes, const ConvShape &s, View x, View mask, const bf16_t *w,
                   const float *b, View y, int relu, hipStream_t st) {
  if (lds) {
    if constexpr (NB % 2 == 0 && G == 2) {  // several output blocks: the wavefronts split them
      constexpr int WS = NB % 4 == 0 ? 4 : 2;
      // four k-steps of operands in flight where they pay (a wavefront
      // loads only its NB / WS weight fragments a k-step): at 256x512 the
      // 2-, 4- and 8-block layers 48.2 -> 45.8, 19.1 -> 18.4 and
      // 20.5 -> 19.0 us; the 6-block layer (three fragments a wavefront)
      // 35.3 -> 40.0 us, so it keeps two
      constexpr int GN = NB == 2 || NB == 4 || NB == 8 ? 4 : 2;
      if (bytes > 65536)
        SCIHPC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_conv_ldsn<NB, DIM, WS, GN>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kConvLdsMax));
      k_conv_ldsn<NB, DIM, WS, GN><<<grid, kConvThreads, bytes, st>>>(x, mask, w, b, y, s, relu);
      return;
    }
    if (bytes > 65536)  // past the default dynamic LDS limit (a host-side attribute, cheap to set again)
      SCIHPC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_conv_lds<NB, DIM, G>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kConvLdsMax));
    k_conv_lds<NB, DIM, G><<<grid, kConvThreads, bytes, st>>>(x, mask, w, b, y, s, relu);
  } else {
    k_conv<NB, DIM, G><<<grid, kConvThreads, 0, st>>>(x, mask, w, b, y, s, relu);
  }
}
// deep reductions (32 k-steps or more) over an LDS tile with one or two
// output blocks keep eight k-steps of weight operands in flight instead of
// two (the 192-channel 64x128 layer 101 -> 91 us; through the caches, at
// 32x64, no gain)
template <int NB, int DIM>
void conv_kernel(dim3 grid, bool lds, size_t bytes, const ConvShape &s, View x, View mask, const bf16_t *w,
                 const float *b, View y, int relu, hipStream_t st) {
  if constexpr (NB <= 2) {
    if (lds && (Taps<DIM>::T * s.cin + 31) / 32 >= 32) {
      conv_kernel_g<NB, DIM, 8>(grid, lds, bytes, s, x, mask, w, b, y, relu, st);
      return;
    }
  }
  conv_kernel_g<NB, DIM, 2>(grid, lds, bytes, s, x, mask, w, b, y, relu, st);
}
template <int DIM>
void launch_conv_dim(const ConvShape &s, View x, View mask, const bf16_t *w, const float *b, View y, int relu,
                     hipStream_t st) {
  const unsigned tiles = (unsigned)((long)s.n * s.d * s.h * ((s.w + 63) / 64));
  const int blocks = s.cout / 16;
  const size_t bytes = conv_lds_bytes(DIM, s.cin);
  bool lds = bytes <= kConvLdsMax;
  if ((long)tiles < kSplitBelow && blocks > 1) {
    const int per = blocks % 2 == 0 && (long)tiles * blocks >= 4 * kSplitBelow ? 2 : 1;
    const dim3 grid(tiles, blocks / per);
    lds = lds && (int)grid.y <= kConvLdsSplitMax;
    if (per == 2) conv_kernel<2, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st);
    else conv_kernel<1, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st);
    SCIHPC_CHECK_LAUNCH();
    return;
  }
  // large, narrow layers: weights in registers across several tiles
  if (DIM == 2 && (long)tiles >= 2048) {
    const int ks = (Taps<DIM>::T * s.cin + 31) / 32;
#define SCIHPC_RW(NB_, KS_)                                                                              \
  if (blocks == NB_ && ks == KS_) {                                                                      \
    const unsigned g = (tiles + rw_tiles<KS_>() - 1) / rw_tiles<KS_>();                                  \
    k_conv_rw<NB_, DIM, KS_><<<g, kConvThreads, 0, st>>>(x, mask, w, b, y, s, relu, (int)tiles);          \
    SCIHPC_CHECK_LAUNCH();                                                                               \
    return;                                                                                              \
  }
    SCIHPC_RW(1, 5)
    SCIHPC_RW(1, 9)
    SCIHPC_RW(1, 14)
    SCIHPC_RW(2, 5)
    SCIHPC_RW(2, 9)
    SCIHPC_RW(3, 5)
#undef SCIHPC_RW
  }
  const dim3 grid(tiles);
  switch (blocks) {
    case 1: conv_kernel<1, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    case 2: conv_kernel<2, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    case 3: conv_kernel<3, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    case 4: conv_kernel<4, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    case 6: conv_kernel<6, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    case 8: conv_kernel<8, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    case 12: conv_kernel<12, DIM>(grid, lds, bytes, s, x, mask, w, b, y, relu, st); break;
    default: throw Error(SCIHPC_BAD_ARG, "unsupported convolution width");
  }
  SCIHPC_CHECK_LAUNCH();
}
inline void launch_conv(int ndim, const ConvShape &s, View x, View mask, const bf16_t *w, const float *b, View y,
                        int relu, hipStream_t st) {
  if (ndim == 3) launch_conv_dim<3>(s, x, mask, w, b, y, relu, st);
  else launch_conv_dim<2>(s, x, mask, w, b, y, relu, st);
}

struct Layer {
  int cin, cout, lev, taps;
  long off;  // w at off, b at off + cout*taps*cin
  long nw() const { return (long)cout * taps * cin; }
  long size() const { return nw() + cout; }
};

}  // namespace
}  // namespace scihpc

using namespace scihpc;

struct scihpc_unet {
  int ndim = 2, T = 9;  // 2-D: 3x3 convolutions; 3-D: 3x3x3
  int D = 1, H = 0, W = 0, nmax = 0;
  hipStream_t stream = nullptr;
  Layer L[kLayers];
  long head_off = 0, np = 0;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  bf16_t *Pb = nullptr, *Pt = nullptr;
  long step = 0;
  float out_scale = 1.f;
  // activations (BF16, NHWC) and their gradients
  bf16_t *in0, *t0a, *cat0, *p1, *t1a, *cat1, *p2, *t2a, *cat2, *p3, *t3a, *t3b, *t2c, *t2d, *t1c, *t1d, *t0c, *t0d;
  bf16_t *g_t0a, *g_cat0, *g_p1, *g_t1a, *g_cat1, *g_p2, *g_t2a, *g_cat2, *g_p3, *g_t3a, *g_t3b, *g_t2c, *g_t2d,
      *g_t1c, *g_t1d, *g_t0c, *g_t0d;
  float *out = nullptr, *tgt = nullptr, *dout = nullptr, *part = nullptr;
  long part_n = 0;
  double *lpart = nullptr, *host_l = nullptr;
  int *d_idx = nullptr;
  float *store_feat = nullptr, *store_tgt = nullptr, *store_face = nullptr, *store_e = nullptr;
  int store_n = 0;
  bool energy = false;  // loss: energy norm of the error (faces given) or mean squared error
  int smooth = 0;       // damped-Jacobi sweeps applied to the error inside the energy loss
  float omega = 0.8f;
  float *sm0 = nullptr, *sm1 = nullptr;  // smoothed errors / their gradients
  float *bface = nullptr, *be = nullptr;  // faces and target energy of the current batch
  int n_fwd = 0;  // batch of the last forward pass
  std::vector<void *> allocs;

  int dz(int lev) const { return ndim == 3 ? D >> lev : 1; }
  long pix(int lev, int n) const { return (long)n * dz(lev) * (H >> lev) * (W >> lev); }

  template <typename T>
  T *alloc(long n) {
    void *p = nullptr;
    SCIHPC_HIP(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    allocs.push_back(p);
    SCIHPC_HIP(hipMemsetAsync(p, 0, sizeof(T) * (size_t)(n > 0 ? n : 1), stream));
    return static_cast<T *>(p);
  }
  bf16_t *act(int lev, int c) { return alloc<bf16_t>(pix(lev, nmax) * c); }

  void create(int nd, int d, int h, int w, int n, unsigned long long seed) {
    SCIHPC_REQUIRE(nd == 2 || nd == 3, "the U-Net is 2-D or 3-D");
    if (nd == 2) d = 1;
    SCIHPC_REQUIRE(h >= 8 && w >= 8 && h % 8 == 0 && w % 8 == 0 && (nd == 2 || (d >= 8 && d % 8 == 0)),
                   "U-Net size must be a multiple of 8, >= 8");
    SCIHPC_REQUIRE(n >= 1 && (long)n * d * h * w <= (1L << 26), "batch x pixels must be in [1, 2^26]");
    ndim = nd;
    T = nd == 3 ? 27 : 9;
    D = d;
    H = h;
    W = w;
    nmax = n;
    SCIHPC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const int spec[kLayers][3] = {{16, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2},
                                  {64, 64, 2},  {64, 128, 3}, {128, 128, 3}, {192, 64, 2}, {64, 64, 2},
                                  {96, 32, 1},  {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};
    long off = 0;
    for (int i = 0; i < kLayers; ++i) {
      L[i] = Layer{spec[i][0], spec[i][1], spec[i][2], T, off};
      off += L[i].size();
    }
    head_off = off;
    np = off + kHeadC + 1;
    P = alloc<float>(np);
    G = alloc<float>(np);
    M = alloc<float>(np);
    V = alloc<float>(np);
    Pb = alloc<bf16_t>(np);
    Pt = alloc<bf16_t>(np);
    // activations
    in0 = act(0, 16); t0a = act(0, 16); cat0 = act(0, 48); p1 = act(1, 16); t1a = act(1, 32); cat1 = act(1, 96);
    p2 = act(2, 32); t2a = act(2, 64); cat2 = act(2, 192); p3 = act(3, 64); t3a = act(3, 128); t3b = act(3, 128);
    t2c = act(2, 64); t2d = act(2, 64); t1c = act(1, 32); t1d = act(1, 32); t0c = act(0, 16); t0d = act(0, 16);
    g_t0a = act(0, 16); g_cat0 = act(0, 48); g_p1 = act(1, 16); g_t1a = act(1, 32); g_cat1 = act(1, 96);
    g_p2 = act(2, 32); g_t2a = act(2, 64); g_cat2 = act(2, 192); g_p3 = act(3, 64); g_t3a = act(3, 128);
    g_t3b = act(3, 128); g_t2c = act(2, 64); g_t2d = act(2, 64); g_t1c = act(1, 32); g_t1d = act(1, 32);
    g_t0c = act(0, 16); g_t0d = act(0, 16);
    out = alloc<float>(pix(0, nmax));
    tgt = alloc<float>(pix(0, nmax));
    dout = alloc<float>(pix(0, nmax));
    sm0 = alloc<float>(pix(0, nmax));
    sm1 = alloc<float>(pix(0, nmax));
    lpart = alloc<double>(nblk(pix(0, nmax)));
    d_idx = alloc<int>(2 * nmax);
    bface = alloc<float>((long)ndim * pix(0, nmax));
    be = alloc<float>(nmax);
    part_n = (long)nblk(pix(0, nmax)) * (kHeadC + 1);
    for (int i = 0; i < kLayers; ++i) {
      int chunks, cpix;
      wgrad_chunks(i, nmax, chunks, cpix);
      part_n = std::max(part_n, (long)chunks * L[i].size());
    }
    part = alloc<float>(part_n);
    SCIHPC_HIP(hipHostMalloc(reinterpret_cast<void **>(&host_l), sizeof(double) * nblk(pix(0, nmax)),
                             hipHostMallocDefault));
    init_params(seed);
  }

  void destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (float *b : {store_feat, store_tgt, store_face, store_e})
      if (b) (void)hipFree(b);
    store_feat = store_tgt = store_face = store_e = nullptr;
    for (void *p : allocs) (void)hipFree(p);
    allocs.clear();
    if (host_l) (void)hipHostFree(host_l);
    host_l = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  // He-normal weights (fan-in T cin; the first layer's fan-in counts its
  // two real input channels), zero biases, a small head
  void init_params(unsigned long long seed) {
    std::vector<float> h(np, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int i = 0; i < kLayers; ++i) {
      const int fan = T * (i == 0 ? 2 : L[i].cin);
      const double sd = std::sqrt(2.0 / fan);
      for (long e = 0; e < L[i].nw(); ++e) {
        const int ci = (int)(e % L[i].cin);
        h[L[i].off + e] = (i == 0 && ci >= 2) ? 0.f : (float)(sd * nd(rng));
      }
    }
    for (int c = 0; c < kHeadC; ++c) h[head_off + c] = (float)(0.25 * nd(rng));
    set_params(h.data());
  }

  void refresh_derived() {
    k_to_bf16<<<nblk(np), kEB, 0, stream>>>(P, Pb, np);
    for (int i = 0; i < kLayers; ++i)
      k_flip<<<nblk(L[i].nw()), kEB, 0, stream>>>(Pb + L[i].off, Pt + L[i].off, L[i].cin, L[i].cout, T);
    SCIHPC_CHECK_LAUNCH();
  }
  void set_params(const float *host) {
    SCIHPC_HIP(hipMemcpyAsync(P, host, sizeof(float) * np, hipMemcpyHostToDevice, stream));
    SCIHPC_HIP(hipMemsetAsync(M, 0, sizeof(float) * np, stream));
    SCIHPC_HIP(hipMemsetAsync(V, 0, sizeof(float) * np, stream));
    step = 0;
    refresh_derived();
    SCIHPC_HIP(hipStreamSynchronize(stream));
  }

  // ------------------------------------------------------- launches
  ConvShape shape(int i, int n) const {
    return ConvShape{n, dz(L[i].lev), H >> L[i].lev, W >> L[i].lev, L[i].cin, L[i].cout};
  }
  static unsigned conv_grid(const ConvShape &s) { return (unsigned)((long)s.n * s.d * s.h * ((s.w + 63) / 64)); }

  void conv_fwd(int i, int n, View x, View y, hipStream_t st) {
    const ConvShape s = shape(i, n);
    launch_conv(ndim, s, x, View{nullptr, 0, 0}, Pb + L[i].off, P + L[i].off + L[i].nw(), y, 1, st);
  }
  // gx = conv^T(gy * (y > 0))
  void conv_bwd_data(int i, int n, View gy, View y, View gx) {
    ConvShape s = shape(i, n);
    std::swap(s.cin, s.cout);
    launch_conv(ndim, s, gy, y, Pt + L[i].off, nullptr, gx, 0, stream);
  }
  void wgrad_chunks(int i, int n, int &chunks, int &cpix) const {
    const long npix = pix(L[i].lev, n);
    const int pairs = (L[i].cout / 16) * (L[i].cin / 16) * (T / 9);
    // about kWgradWGs workgroups, pixel chunks of at least 128
    long c = std::max(1, (kWgradWGs + pairs - 1) / pairs);
    c = std::min(c, std::max(1L, npix / 128));
    long cp = (npix + c - 1) / c;
    cp = (cp + 127) / 128 * 128;
    chunks = (int)((npix + cp - 1) / cp);
    cpix = (int)cp;
  }
  void conv_wgrad(int i, int n, View x, View gy, View y) {
    const ConvShape s = shape(i, n);
    int chunks, cpix;
    wgrad_chunks(i, n, chunks, cpix);
    const unsigned grid = (unsigned)((L[i].cout / 16) * (L[i].cin / 16) * (T / 9) * chunks);
    if (ndim == 3) k_conv_wgrad<3><<<grid, kConvThreads, 0, stream>>>(x, gy, y, part, s, cpix);
    else k_conv_wgrad<2><<<grid, kConvThreads, 0, stream>>>(x, gy, y, part, s, cpix);
    k_sum_rows_2d<<<ceil_div(L[i].size(), kSumCols), dim3(kSumCols, kSumRowGroups), 0, stream>>>(
        part, chunks, L[i].size(), G + L[i].off);
    SCIHPC_CHECK_LAUNCH();
  }
  void conv_back(int i, int n, View x, View y, View gy, View gx) {
    conv_wgrad(i, n, x, gy, y);
    if (gx.p) conv_bwd_data(i, n, gy, y, gx);
  }
  void pool(View src, View dst, int lev, int n, int c, hipStream_t st) {  // lev = coarse level
    const int d = dz(lev), h = H >> lev, w = W >> lev;
    const unsigned g = nblk((long)n * d * h * w * (c / kVec));
    if (ndim == 3) k_pool2<3><<<g, kEB, 0, st>>>(src, dst, n, d, h, w, c);
    else k_pool2<2><<<g, kEB, 0, st>>>(src, dst, n, d, h, w, c);
    SCIHPC_CHECK_LAUNCH();
  }
  void up(View src, View dst, int lev, int n, int c, hipStream_t st) {  // lev = coarse level
    const int d = dz(lev), h = H >> lev, w = W >> lev;
    const unsigned g = nblk(pix(lev - 1, n) * (c / kVec));
    if (ndim == 3) k_up2<3><<<g, kEB, 0, st>>>(src, dst, n, d, h, w, c);
    else k_up2<2><<<g, kEB, 0, st>>>(src, dst, n, d, h, w, c);
    SCIHPC_CHECK_LAUNCH();
  }
  void up_bwd(View gf, View gc, int lev, int n, int c) {  // lev = coarse level
    const int d = dz(lev), h = H >> lev, w = W >> lev;
    const unsigned g = nblk(pix(lev, n) * (c / kVec));
    if (ndim == 3) k_up2_bwd<3><<<g, kEB, 0, stream>>>(gf, gc, n, d, h, w, c);
    else k_up2_bwd<2><<<g, kEB, 0, stream>>>(gf, gc, n, d, h, w, c);
    SCIHPC_CHECK_LAUNCH();
  }
  void pool_bwd_add(View gf, View gc, int lev, int n, int c) {  // lev = coarse level
    const int d = dz(lev), h = H >> lev, w = W >> lev;
    const unsigned g = nblk(pix(lev - 1, n) * (c / kVec));
    if (ndim == 3) k_pool2_bwd_add<3><<<g, kEB, 0, stream>>>(gf, gc, n, d, h, w, c);
    else k_pool2_bwd_add<2><<<g, kEB, 0, stream>>>(gf, gc, n, d, h, w, c);
    SCIHPC_CHECK_LAUNCH();
  }

  static View v(bf16_t *p, int stride, int off = 0) { return View{p, stride, off}; }

  void forward(int n, hipStream_t st) {
    SCIHPC_REQUIRE(n >= 1 && n <= nmax, "batch exceeds the U-Net's capacity");
    conv_fwd(0, n, v(in0, 16), v(t0a, 16), st);
    conv_fwd(1, n, v(t0a, 16), v(cat0, 48, 32), st);
    pool(v(cat0, 48, 32), v(p1, 16), 1, n, 16, st);
    conv_fwd(2, n, v(p1, 16), v(t1a, 32), st);
    conv_fwd(3, n, v(t1a, 32), v(cat1, 96, 64), st);
    pool(v(cat1, 96, 64), v(p2, 32), 2, n, 32, st);
    conv_fwd(4, n, v(p2, 32), v(t2a, 64), st);
    conv_fwd(5, n, v(t2a, 64), v(cat2, 192, 128), st);
    pool(v(cat2, 192, 128), v(p3, 64), 3, n, 64, st);
    conv_fwd(6, n, v(p3, 64), v(t3a, 128), st);
    conv_fwd(7, n, v(t3a, 128), v(t3b, 128), st);
    up(v(t3b, 128), v(cat2, 192, 0), 3, n, 128, st);
    conv_fwd(8, n, v(cat2, 192), v(t2c, 64), st);
    conv_fwd(9, n, v(t2c, 64), v(t2d, 64), st);
    up(v(t2d, 64), v(cat1, 96, 0), 2, n, 64, st);
    conv_fwd(10, n, v(cat1, 96), v(t1c, 32), st);
    conv_fwd(11, n, v(t1c, 32), v(t1d, 32), st);
    up(v(t1d, 32), v(cat0, 48, 0), 1, n, 32, st);
    conv_fwd(12, n, v(cat0, 48), v(t0c, 16), st);
    conv_fwd(13, n, v(t0c, 16), v(t0d, 16), st);
    k_head_fwd<<<nblk(pix(0, n)), kEB, 0, st>>>(v(t0d, 16), P + head_off, out, pix(0, n));
    SCIHPC_CHECK_LAUNCH();
    n_fwd = n;
  }

  // loss and all parameter gradients of the last forward pass (into G)
  double backward(bool want_loss) {
    const int n = n_fwd;
    SCIHPC_REQUIRE(n >= 1, "backward needs a forward pass with targets");
    const long np0 = pix(0, n);
    const unsigned nb0 = nblk(np0);
    if (energy)
      energy_loss(n, nb0);
    else
      k_mse<<<nb0, kEB, 0, stream>>>(out, tgt, dout, np0, lpart);
    k_head_bwd<<<nb0, kEB, 0, stream>>>(dout, v(t0d, 16), P + head_off, v(g_t0d, 16), part, np0);
    k_sum_rows_tree<<<kHeadC + 1, kEB, 0, stream>>>(part, (int)nb0, kHeadC + 1, G + head_off);
    SCIHPC_CHECK_LAUNCH();
    const View none{nullptr, 0, 0};
    conv_back(13, n, v(t0c, 16), v(t0d, 16), v(g_t0d, 16), v(g_t0c, 16));
    conv_back(12, n, v(cat0, 48), v(t0c, 16), v(g_t0c, 16), v(g_cat0, 48));
    up_bwd(v(g_cat0, 48, 0), v(g_t1d, 32), 1, n, 32);
    conv_back(11, n, v(t1c, 32), v(t1d, 32), v(g_t1d, 32), v(g_t1c, 32));
    conv_back(10, n, v(cat1, 96), v(t1c, 32), v(g_t1c, 32), v(g_cat1, 96));
    up_bwd(v(g_cat1, 96, 0), v(g_t2d, 64), 2, n, 64);
    conv_back(9, n, v(t2c, 64), v(t2d, 64), v(g_t2d, 64), v(g_t2c, 64));
    conv_back(8, n, v(cat2, 192), v(t2c, 64), v(g_t2c, 64), v(g_cat2, 192));
    up_bwd(v(g_cat2, 192, 0), v(g_t3b, 128), 3, n, 128);
    conv_back(7, n, v(t3a, 128), v(t3b, 128), v(g_t3b, 128), v(g_t3a, 128));
    conv_back(6, n, v(p3, 64), v(t3a, 128), v(g_t3a, 128), v(g_p3, 64));
    pool_bwd_add(v(g_cat2, 192, 128), v(g_p3, 64), 3, n, 64);
    conv_back(5, n, v(t2a, 64), v(cat2, 192, 128), v(g_cat2, 192, 128), v(g_t2a, 64));
    conv_back(4, n, v(p2, 32), v(t2a, 64), v(g_t2a, 64), v(g_p2, 32));
    pool_bwd_add(v(g_cat1, 96, 64), v(g_p2, 32), 2, n, 32);
    conv_back(3, n, v(t1a, 32), v(cat1, 96, 64), v(g_cat1, 96, 64), v(g_t1a, 32));
    conv_back(2, n, v(p1, 16), v(t1a, 32), v(g_t1a, 32), v(g_p1, 16));
    pool_bwd_add(v(g_cat0, 48, 32), v(g_p1, 16), 1, n, 16);
    conv_back(1, n, v(t0a, 16), v(cat0, 48, 32), v(g_cat0, 48, 32), v(g_t0a, 16));
    conv_back(0, n, v(in0, 16), v(t0a, 16), v(g_t0a, 16), none);
    SCIHPC_CHECK_LAUNCH();
    if (!want_loss) return 0.0;
    SCIHPC_HIP(hipMemcpyAsync(host_l, lpart, sizeof(double) * nb0, hipMemcpyDeviceToHost, stream));
    SCIHPC_HIP(hipStreamSynchronize(stream));
    double s = 0.0;
    for (unsigned b = 0; b < nb0; ++b) s += host_l[b];
    return s;
  }

  // energy loss of the last forward pass into lpart and its gradient into
  // dout, through `smooth` Jacobi sweeps of the error
  void energy_loss(int n, unsigned nb0) {
    const int d = dz(0);
    auto jac = [&](const float *a, float *b, int tr) {
      if (ndim == 3) k_jacobi<3><<<nb0, kEB, 0, stream>>>(a, b, bface, n, d, H, W, omega, tr);
      else k_jacobi<2><<<nb0, kEB, 0, stream>>>(a, b, bface, n, d, H, W, omega, tr);
    };
    auto loss = [&](const float *o, const float *t, float *g) {
      if (ndim == 3) k_energy_loss<3><<<nb0, kEB, 0, stream>>>(o, t, bface, be, g, n, d, H, W, lpart);
      else k_energy_loss<2><<<nb0, kEB, 0, stream>>>(o, t, bface, be, g, n, d, H, W, lpart);
    };
    if (smooth == 0) {
      loss(out, tgt, dout);
      SCIHPC_CHECK_LAUNCH();
      return;
    }
    // e_0 = out - tgt, e_{k+1} = M e_k (ping-pong through sm0 / sm1)
    k_sub<<<nb0, kEB, 0, stream>>>(out, tgt, sm0, pix(0, n));
    float *cur = sm0, *nxt = sm1;
    for (int k = 0; k < smooth; ++k) {
      jac(cur, nxt, 0);
      std::swap(cur, nxt);
    }
    // g_S = dE(e_S)/de_S, g_k = M^T g_{k+1}; the last one lands in dout
    loss(cur, nullptr, nxt);
    std::swap(cur, nxt);  // cur = g_S
    for (int k = 0; k < smooth; ++k) {
      float *dst = k + 1 == smooth ? dout : (cur == sm0 ? sm1 : sm0);
      jac(cur, dst, 1);
      cur = dst;
    }
    SCIHPC_CHECK_LAUNCH();
  }

  void adam(float lr) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - (float)std::pow((double)b1, (double)step);
    const float c2 = 1.f - (float)std::pow((double)b2, (double)step);
    k_adam<<<nblk(np), kEB, 0, stream>>>(P, G, M, V, Pb, np, lr, b1, b2, 1e-8f, c1, c2);
    SCIHPC_CHECK_LAUNCH();
    for (int i = 0; i < kLayers; ++i)
      k_flip<<<nblk(L[i].nw()), kEB, 0, stream>>>(Pb + L[i].off, Pt + L[i].off, L[i].cin, L[i].cout, T);
    SCIHPC_CHECK_LAUNCH();
  }

  void make_input(const float *feat, const float *tstore, const float *fstore, const float *estore, const int *ix,
                  float *t, float *f, int n, float inv_scale, hipStream_t st) {
    if (ndim == 3)
      k_make_input<3><<<nblk(pix(0, n)), kEB, 0, st>>>(feat, tstore, fstore, estore, ix, in0, t, f, be, n, D, H, W,
                                                        inv_scale);
    else
      k_make_input<2><<<nblk(pix(0, n)), kEB, 0, st>>>(feat, tstore, fstore, estore, ix, in0, t, f, be, n, 1, H, W,
                                                        inv_scale);
  }
  // batch from the device sample store; idx = host [n][2] (sample, mirror bits)
  void load_batch(int n, const int *idx) {
    SCIHPC_REQUIRE(store_feat && store_n > 0, "no training samples stored");
    SCIHPC_REQUIRE(n >= 1 && n <= nmax, "batch exceeds the U-Net's capacity");
    for (int b = 0; b < n; ++b)
      SCIHPC_REQUIRE(idx[2 * b] >= 0 && idx[2 * b] < store_n && idx[2 * b + 1] >= 0 &&
                         idx[2 * b + 1] < (1 << ndim),
                     "bad sample index or flip code");
    SCIHPC_HIP(hipMemcpyAsync(d_idx, idx, sizeof(int) * 2 * n, hipMemcpyHostToDevice, stream));
    make_input(store_feat, store_tgt, energy ? store_face : nullptr, store_e, d_idx, tgt,
               energy ? bface : nullptr, n, 1.f / out_scale, stream);
    SCIHPC_CHECK_LAUNCH();
  }
};

// ------------------------------------------------ multigrid hook
namespace scihpc {
void unet_infer(scihpc_unet *net, const float *feat, float *out, hipStream_t stream) {
  net->make_input(feat, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1, 1.f, stream);
  SCIHPC_CHECK_LAUNCH();
  net->forward(1, stream);
  SCIHPC_HIP(hipMemcpyAsync(out, net->out, sizeof(float) * net->pix(0, 1), hipMemcpyDeviceToDevice, stream));
}
int unet_ndim(const scihpc_unet *net) { return net->ndim; }
int unet_depth(const scihpc_unet *net) { return net->D; }
int unet_height(const scihpc_unet *net) { return net->H; }
int unet_width(const scihpc_unet *net) { return net->W; }
float unet_out_scale(const scihpc_unet *net) { return net->out_scale; }
}  // namespace scihpc

// ================================================================ C ABI
namespace {
int unet_create(int ndim, int depth, int height, int width, int max_batch, unsigned long long seed,
                scihpc_unet **out) {
  return guarded([&] {
    SCIHPC_REQUIRE(out, "null output handle");
    *out = nullptr;
    auto *h = new scihpc_unet();
    try {
      h->create(ndim, depth, height, width, max_batch, seed);
    } catch (...) {
      h->destroy();
      delete h;
      throw;
    }
    *out = h;
  });
}
}  // namespace

extern "C" int scihpc_unet_create(int height, int width, int max_batch, unsigned long long seed, scihpc_unet **out) {
  return unet_create(2, 1, height, width, max_batch, seed, out);
}

extern "C" int scihpc_unet_create3(int depth, int height, int width, int max_batch, unsigned long long seed,
                                   scihpc_unet **out) {
  return unet_create(3, depth, height, width, max_batch, seed, out);
}

extern "C" int scihpc_unet_destroy(scihpc_unet *h) {
  return guarded([&] {
    if (!h) return;
    h->destroy();
    delete h;
  });
}

extern "C" int scihpc_unet_num_params(scihpc_unet *h, long long *n) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && n, "null argument");
    *n = h->np;
  });
}

extern "C" int scihpc_unet_layer(scihpc_unet *h, int layer, int *cin, int *cout, int *level, long long *offset) {
  return guarded([&] {
    SCIHPC_REQUIRE(h, "null handle");
    SCIHPC_REQUIRE(layer >= 0 && layer <= kLayers, "layer out of range");
    if (layer == kLayers) {  // the 1x1 head: 16 weights then 1 bias
      if (cin) *cin = kHeadC;
      if (cout) *cout = 1;
      if (level) *level = 0;
      if (offset) *offset = h->head_off;
      return;
    }
    if (cin) *cin = h->L[layer].cin;
    if (cout) *cout = h->L[layer].cout;
    if (level) *level = h->L[layer].lev;
    if (offset) *offset = h->L[layer].off;
  });
}

extern "C" int scihpc_unet_get_params(scihpc_unet *h, float *host) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && host, "null argument");
    SCIHPC_HIP(hipMemcpyAsync(host, h->P, sizeof(float) * h->np, hipMemcpyDeviceToHost, h->stream));
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_unet_set_params(scihpc_unet *h, const float *host) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && host, "null argument");
    h->set_params(host);
  });
}

extern "C" int scihpc_unet_get_grads(scihpc_unet *h, float *host) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && host, "null argument");
    SCIHPC_HIP(hipMemcpyAsync(host, h->G, sizeof(float) * h->np, hipMemcpyDeviceToHost, h->stream));
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

extern "C" int scihpc_unet_set_out_scale(scihpc_unet *h, float s) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && std::isfinite(s) && s > 0.f, "output scale must be finite and > 0");
    h->out_scale = s;
  });
}

extern "C" int scihpc_unet_get_out_scale(scihpc_unet *h, float *s) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && s, "null argument");
    *s = h->out_scale;
  });
}

// features: host [n][pixels][2] FP32; targets: host [n][pixels]; faces:
// host [n][pixels][ndim] (NULL: mean-squared-error loss, else energy-norm
// loss); pixels = (D x) H x W, x fastest
extern "C" int scihpc_unet_set_samples(scihpc_unet *h, int n, const float *feat, const float *target,
                                       const float *faces) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && feat && target && n >= 1, "bad sample arguments");
    const long hw = h->pix(0, 1);
    const int nd = h->ndim;
    std::vector<float> e;
    if (faces) {
      // energy norm of each target (mirror invariant), FP64 on the host
      e.resize(n);
      const int cnt[3] = {h->W, h->H, h->dz(0)};
      const long step[3] = {1, (long)h->W, (long)h->W * h->H};
      for (int s = 0; s < n; ++s) {
        double acc = 0.0;
        const float *t = target + (long)s * hw;
        const float *f = faces + (long)s * hw * nd;
        for (long p = 0; p < hw; ++p) {
          const int pos[3] = {(int)(p % h->W), (int)((p / h->W) % h->H), (int)(p / ((long)h->W * h->H))};
          for (int a = 0; a < nd; ++a)
            if (pos[a] + 1 < cnt[a]) {
              const double dd = (double)t[p] - t[p + step[a]];
              acc += (double)f[nd * p + a] * dd * dd;
            }
        }
        SCIHPC_REQUIRE(acc > 0.0 && std::isfinite(acc), "a target has zero or non-finite energy");
        e[s] = (float)(acc / ((double)h->out_scale * h->out_scale));
      }
    }
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
    for (float **b : {&h->store_feat, &h->store_tgt, &h->store_face, &h->store_e})
      if (*b) {
        SCIHPC_HIP(hipFree(*b));
        *b = nullptr;
      }
    h->store_n = 0;
    SCIHPC_HIP(hipMalloc(reinterpret_cast<void **>(&h->store_feat), sizeof(float) * 2 * hw * n));
    SCIHPC_HIP(hipMalloc(reinterpret_cast<void **>(&h->store_tgt), sizeof(float) * hw * n));
    SCIHPC_HIP(hipMemcpyAsync(h->store_feat, feat, sizeof(float) * 2 * hw * n, hipMemcpyHostToDevice, h->stream));
    SCIHPC_HIP(hipMemcpyAsync(h->store_tgt, target, sizeof(float) * hw * n, hipMemcpyHostToDevice, h->stream));
    if (faces) {
      SCIHPC_HIP(hipMalloc(reinterpret_cast<void **>(&h->store_face), sizeof(float) * nd * hw * n));
      SCIHPC_HIP(hipMalloc(reinterpret_cast<void **>(&h->store_e), sizeof(float) * n));
      SCIHPC_HIP(
          hipMemcpyAsync(h->store_face, faces, sizeof(float) * nd * hw * n, hipMemcpyHostToDevice, h->stream));
      SCIHPC_HIP(hipMemcpyAsync(h->store_e, e.data(), sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    }
    h->energy = faces != nullptr;
    h->store_n = n;
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

// forward on stored samples idx[n][2] = (sample, mirror bits: 1 = x, 2 = y,
// 4 = z); out (host [n][pixels], NULL to skip) receives the network output
extern "C" int scihpc_unet_forward(scihpc_unet *h, int n, const int *idx, float *out) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && idx, "null argument");
    h->load_batch(n, idx);
    h->forward(n, h->stream);
    if (out) {
      SCIHPC_HIP(hipMemcpyAsync(out, h->out, sizeof(float) * h->pix(0, n), hipMemcpyDeviceToHost, h->stream));
    }
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

// loss (energy norm or mean squared error of the error against the stored
// targets / out_scale) and gradients of the last forward pass
extern "C" int scihpc_unet_backward(scihpc_unet *h, double *loss) {
  return guarded([&] {
    SCIHPC_REQUIRE(h, "null handle");
    const double l = h->backward(true);
    if (loss) *loss = l;
  });
}

// one Adam step: forward + backward + update; loss (optional) is the batch
// loss before the update
extern "C" int scihpc_unet_train_step(scihpc_unet *h, int n, const int *idx, float lr, double *loss) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && idx && std::isfinite(lr) && lr > 0.f, "bad training arguments");
    h->load_batch(n, idx);
    h->forward(n, h->stream);
    const double l = h->backward(loss != nullptr);
    h->adam(lr);
    if (loss) *loss = l;
  });
}

// energy loss through `sweeps` damped-Jacobi sweeps (weight omega) of the
// error on the network's grid (0: the plain energy norm)
extern "C" int scihpc_unet_set_loss_smoothing(scihpc_unet *h, int sweeps, float omega) {
  return guarded([&] {
    SCIHPC_REQUIRE(h, "null handle");
    SCIHPC_REQUIRE(sweeps >= 0 && sweeps <= 16, "sweeps must be in [0, 16]");
    SCIHPC_REQUIRE(std::isfinite(omega) && omega > 0.f && omega <= 1.f, "omega must be in (0, 1]");
    h->smooth = sweeps;
    h->omega = omega;
  });
}

extern "C" int scihpc_unet_sync(scihpc_unet *h) {
  return guarded([&] {
    SCIHPC_REQUIRE(h, "null handle");
    SCIHPC_HIP(hipStreamSynchronize(h->stream));
  });
}

// device time of `iters` forward passes (which = 0) or training steps
// (which = 1: forward, backward, Adam) on stored samples 0..n-1 (HIP events
// on the network's stream), and the matrix-core FLOPs of one: 2 x T x cin x
// cout per pixel for every convolution of T taps, tripled in training for
// the data and weight gradients (the first layer has no data gradient)
extern "C" int scihpc_unet_time(scihpc_unet *h, int n, int which, int iters, double *ms, double *flops) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && iters > 0 && (which == 0 || which == 1), "bad timing request");
    SCIHPC_REQUIRE(h->store_n >= 1, "store samples first");
    std::vector<int> idx(2 * n);
    for (int b = 0; b < n; ++b) {
      idx[2 * b] = b % h->store_n;
      idx[2 * b + 1] = 0;
    }
    double f = 0.0;
    for (int i = 0; i < kLayers; ++i) {
      const double c = 2.0 * h->T * h->L[i].cin * h->L[i].cout * (double)h->pix(h->L[i].lev, n);
      f += which == 0 ? c : (i == 0 ? 2.0 * c : 3.0 * c);
    }
    h->load_batch(n, idx.data());
    auto once = [&] {
      h->forward(n, h->stream);
      if (which == 1) {
        h->backward(false);
        h->adam(1e-12f);
      }
    };
    once();
    hipEvent_t e0, e1;
    SCIHPC_HIP(hipEventCreate(&e0));
    SCIHPC_HIP(hipEventCreate(&e1));
    SCIHPC_HIP(hipEventRecord(e0, h->stream));
    for (int it = 0; it < iters; ++it) once();
    SCIHPC_HIP(hipEventRecord(e1, h->stream));
    SCIHPC_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    SCIHPC_HIP(hipEventElapsedTime(&t, e0, e1));
    SCIHPC_HIP(hipEventDestroy(e0));
    SCIHPC_HIP(hipEventDestroy(e1));
    if (ms) *ms = t / iters;
    if (flops) *flops = f;
  });
}

// ------------------------------------------------ single-kernel entry points
// (host arrays, copied to and from the device; used by the kernel tests)
namespace {
int conv_forward(int ndim, const uint16_t *x, int n, int depth, int height, int width, int cin, const uint16_t *w,
                 const float *b, int cout, int relu, uint16_t *y) {
  return guarded([&] {
    SCIHPC_REQUIRE(x && w && y && n > 0 && depth > 0 && height > 0 && width > 0, "bad convolution arguments");
    SCIHPC_REQUIRE(cin % 16 == 0 && cin >= 16 && cin <= 256, "cin must be a multiple of 16 in [16, 256]");
    const int nb = cout / 16;
    SCIHPC_REQUIRE(cout % 16 == 0 && (nb == 1 || nb == 2 || nb == 3 || nb == 4 || nb == 6 || nb == 8 || nb == 12),
                   "cout must be 16 x {1,2,3,4,6,8,12}");
    const long npix = (long)n * depth * height * width;
    SCIHPC_REQUIRE(npix <= (1L << 26), "too many pixels");
    const int T = ndim == 3 ? 27 : 9;
    const size_t sx = sizeof(bf16_t) * npix * cin, sy = sizeof(bf16_t) * npix * cout;
    const size_t sw = sizeof(bf16_t) * (size_t)cout * T * cin, sb = sizeof(float) * cout;
    hipStream_t st = nullptr;
    void *dx = nullptr, *dy = nullptr, *dw = nullptr, *db = nullptr;
    auto cleanup = [&] {
      if (st) (void)hipStreamSynchronize(st);
      for (void *p : {dx, dy, dw, db})
        if (p) (void)hipFree(p);
      if (st) (void)hipStreamDestroy(st);
    };
    try {
      SCIHPC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      SCIHPC_HIP(hipMalloc(&dx, sx));
      SCIHPC_HIP(hipMalloc(&dy, sy));
      SCIHPC_HIP(hipMalloc(&dw, sw));
      SCIHPC_HIP(hipMalloc(&db, sb));
      SCIHPC_HIP(hipMemcpyAsync(dx, x, sx, hipMemcpyHostToDevice, st));
      SCIHPC_HIP(hipMemcpyAsync(dw, w, sw, hipMemcpyHostToDevice, st));
      if (b) SCIHPC_HIP(hipMemcpyAsync(db, b, sb, hipMemcpyHostToDevice, st));
      const ConvShape s{n, depth, height, width, cin, cout};
      const View xv{static_cast<bf16_t *>(dx), cin, 0}, yv{static_cast<bf16_t *>(dy), cout, 0}, none{nullptr, 0, 0};
      launch_conv(ndim, s, xv, none, static_cast<const bf16_t *>(dw), b ? static_cast<const float *>(db) : nullptr,
                  yv, relu, st);
      SCIHPC_HIP(hipMemcpyAsync(y, dy, sy, hipMemcpyDeviceToHost, st));
      SCIHPC_HIP(hipStreamSynchronize(st));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}
}  // namespace

extern "C" int scihpc_conv3x3_forward(const uint16_t *x, int n, int height, int width, int cin, const uint16_t *w,
                                      const float *b, int cout, int relu, uint16_t *y) {
  return conv_forward(2, x, n, 1, height, width, cin, w, b, cout, relu, y);
}

extern "C" int scihpc_conv3x3x3_forward(const uint16_t *x, int n, int depth, int height, int width, int cin,
                                        const uint16_t *w, const float *b, int cout, int relu, uint16_t *y) {
  return conv_forward(3, x, n, depth, height, width, cin, w, b, cout, relu, y);
}
