// unet_kernels.hpp -- BF16 convolution kernels of the learned pressure
// preconditioner for gfx950.
//
// Activations are NHWC BF16 tensors addressed through channel views (base
// pointer, channel stride of the underlying tensor, channel offset), so a
// skip connection is a slice of the decoder's concatenation buffer and no
// concatenation copy exists.  The 3x3 convolution is an implicit GEMM on
// the 16x16x16 BF16 matrix cores with FP32 accumulation:
//   M = output pixels (a workgroup owns 64 pixels of one image row, each
//       wavefront 16 of them), N = output channels (16-wide blocks, all held
//       by the wavefront), K = 9 taps x input channels (16-channel chunks).
// Operand layout of v_mfma_f32_16x16x16bf16_1k (lane l of 64):
//   A (16 x 16, M x K): row l % 16, k = 4 (l / 16) .. +3
//   B (16 x 16, K x N): k = 4 (l / 16) .. +3, column l % 16
//   C (16 x 16, float4): rows 4 (l / 16) + j, column l % 16, j = 0..3
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scihpc {

typedef uint16_t bf16_t;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ float bf16_to_f(bf16_t b) {
  const unsigned u = ((unsigned)b) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
// round to nearest even (finite inputs)
__host__ __device__ __forceinline__ bf16_t f_to_bf16(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

struct View {  // NHWC channel slice
  bf16_t *p;
  int stride;  // channels of the underlying tensor
  int off;     // first channel of the slice
};

struct ConvShape {
  int n, h, w;     // images, rows, columns
  int cin, cout;   // channels (cin multiple of 16, cout multiple of 16)
};

constexpr int kConvThreads = 256;  // 4 wavefronts x 16 pixels = 64 pixels per workgroup

// y = act(conv3x3(x, w) + b); w layout [cout][9][cin] (BF16), b FP32 (may be
// null).  mask: when given, x is read as x * (mask > 0) (ReLU derivative of
// the layer whose output gradient x is, for the backward-data use).
template <int NB>
__global__ void __launch_bounds__(kConvThreads)
    k_conv3x3(View x, View mask, const bf16_t *__restrict__ w, const float *__restrict__ b, View y, ConvShape s,
              int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int segs = (s.w + 63) / 64;
  int blk = blockIdx.x;
  const int seg = blk % segs;
  blk /= segs;
  const int row = blk % s.h, img = blk / s.h;
  const int px = seg * 64 + wave * 16 + (lane & 15);  // pixel of this lane's A row
  const int kq = 4 * (lane >> 4);                    // k offset of this lane inside a chunk
  floatx4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int chunks = s.cin / 16;
  const int K = 9 * chunks;  // k steps of 16: tap-major, then input-channel chunk
  // the operands of G consecutive k steps are loaded before their MFMAs, so
  // G loads are in flight at once instead of one dependent load per step
  constexpr int G = 4;
  for (int k0 = 0; k0 < K; k0 += G) {
    bf16x4 a[G], bw[G][NB];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int kk = k0 + g;
      a[g] = bf16x4{0, 0, 0, 0};
      if (kk < K) {
        const int t = kk / chunks, cc = kk - t * chunks;
        const int yy = row + t / 3 - 1, xx = px + t % 3 - 1;
        if (yy >= 0 && yy < s.h && xx >= 0 && xx < s.w) {
          const long pix = ((long)img * s.h + yy) * s.w + xx;
          a[g] = *reinterpret_cast<const bf16x4 *>(x.p + pix * x.stride + x.off + cc * 16 + kq);
          if (mask.p) {
            const bf16_t *m = mask.p + pix * mask.stride + mask.off + cc * 16 + kq;
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (!(bf16_to_f(m[q]) > 0.f)) a[g][q] = 0;
          }
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          bw[g][nb] =
              *reinterpret_cast<const bf16x4 *>(w + ((long)(nb * 16 + (lane & 15)) * 9 + t) * s.cin + cc * 16 + kq);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (k0 + g < K) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[g], bw[g][nb], acc[nb], 0, 0, 0);
      }
    }
  }
  // epilogue: rows 4 (lane/16) + j of the wave's 16 pixels, column lane % 16
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = nb * 16 + (lane & 15);
    const float bias = b ? b[co] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = seg * 64 + wave * 16 + kq + j;
      if (p >= s.w) continue;
      float v = acc[nb][j] + bias;
      if (relu) v = v > 0.f ? v : 0.f;
      const long pix = ((long)img * s.h + row) * s.w + p;
      y.p[pix * y.stride + y.off + co] = f_to_bf16(v);
    }
  }
}

// Weight and bias gradient of the 3x3 convolution y = relu(conv(x, w) + b):
//   dW[co][t][ci] = sum_p gz[p][co] x[p + off(t)][ci],  db[co] = sum_p gz[p][co]
// with gz = g * (mask > 0) (mask = the layer's stored output).  A = gz^T
// (output channels x pixels), B = x shifted by the tap (pixels x input
// channels).  A workgroup owns one 16 x 16 channel block and one pixel chunk;
// its four wavefronts take interleaved groups of 16 pixels and are summed
// through LDS, and the workgroup writes its partial (no atomics) to
//   part[chunk][co][t][ci], part[chunk][cout*9*cin + co]   (bias, cc == 0 only)
// which k_sum_rows reduces over chunks in a fixed order.
__global__ void __launch_bounds__(kConvThreads)
    k_conv3x3_wgrad(View x, View g, View mask, float *__restrict__ part, ConvShape s, int chunk_pix) {
  __shared__ float red[4][36][64];
  __shared__ float bred[4][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cbs = s.cout / 16, ccs = s.cin / 16;
  int blk = blockIdx.x;
  const int cb = blk % cbs;
  blk /= cbs;
  const int cc = blk % ccs;
  const int chunk = blk / ccs;
  const long npix = (long)s.n * s.h * s.w;
  const long p_begin = (long)chunk * chunk_pix;
  const long p_end = p_begin + chunk_pix < npix ? p_begin + chunk_pix : npix;
  floatx4 acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int co = cb * 16 + (lane & 15), ci = cc * 16 + (lane & 15);
  const int kq = 4 * (lane >> 4);
  for (long p0 = p_begin + wave * 16; p0 < p_end; p0 += 64) {
    bf16x4 a;
    int pr[4], pc[4], pi[4];
    bool pv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long p = p0 + kq + q;
      pv[q] = p < p_end;
      const long pp = pv[q] ? p : p_begin;
      pc[q] = (int)(pp % s.w);
      const long rest = pp / s.w;
      pr[q] = (int)(rest % s.h);
      pi[q] = (int)(rest / s.h);
      float gv = 0.f;
      if (pv[q]) {
        gv = bf16_to_f(g.p[pp * g.stride + g.off + co]);
        if (mask.p && !(bf16_to_f(mask.p[pp * mask.stride + mask.off + co]) > 0.f)) gv = 0.f;
      }
      a[q] = (short)f_to_bf16(gv);  // exact: gv is a BF16 value or zero
      bsum += gv;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      bf16x4 bb;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int yy = pr[q] + t / 3 - 1, xx = pc[q] + t % 3 - 1;
        bf16_t v = 0;
        if (pv[q] && yy >= 0 && yy < s.h && xx >= 0 && xx < s.w)
          v = x.p[(((long)pi[q] * s.h + yy) * s.w + xx) * x.stride + x.off + ci];
        bb[q] = (short)v;
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) red[wave][t * 4 + j][lane] = acc[t][j];
  // lanes l, l^16, l^32, l^48 share an output channel
  bsum += __shfl_xor(bsum, 16, 64);
  bsum += __shfl_xor(bsum, 32, 64);
  if (lane < 16) bred[wave][lane] = bsum;
  __syncthreads();
  const long row = (long)s.cout * 9 * s.cin + s.cout;
  float *dst = part + (long)chunk * row;
  for (int e = threadIdx.x; e < 36 * 64; e += kConvThreads) {
    const int tj = e >> 6, l = e & 63;
    const int t = tj >> 2, j = tj & 3;
    const float v = red[0][tj][l] + red[1][tj][l] + red[2][tj][l] + red[3][tj][l];
    const int o = cb * 16 + 4 * (l >> 4) + j, c = cc * 16 + (l & 15);
    dst[((long)o * 9 + t) * s.cin + c] = v;
  }
  if (cc == 0 && threadIdx.x < 16)
    dst[(long)s.cout * 9 * s.cin + cb * 16 + threadIdx.x] =
        bred[0][threadIdx.x] + bred[1][threadIdx.x] + bred[2][threadIdx.x] + bred[3][threadIdx.x];
}

}  // namespace scihpc
