// unet_kernels.hpp -- BF16 convolution kernels of the learned pressure
// preconditioner for gfx950, 2-D (3x3, 9 taps) and 3-D (3x3x3, 27 taps).
//
// Activations are N(D)HWC BF16 tensors addressed through channel views
// (base pointer, channel stride of the underlying tensor, channel offset),
// so a skip connection is a slice of the decoder's concatenation buffer and
// no concatenation copy exists.  The convolution is an implicit GEMM on
// gfx950's 16x16x32 BF16 matrix cores with FP32 accumulation:
//   M = output pixels (a workgroup owns 64 pixels of one image row, each
//       wavefront 16 of them), N = output channels (16-wide blocks, all held
//       by the wavefront), K = taps x input channels, tap-major, walked in
//       steps of 32 (the last step zero-padded).
// Operand layout of v_mfma_f32_16x16x32_bf16 (lane l of 64):
//   A (16 x 32, M x K): row l % 16, k = 8 (l / 16) .. +7
//   B (32 x 16, K x N): k = 8 (l / 16) .. +7, column l % 16
//   C (16 x 16, float4): rows 4 (l / 16) + j, column l % 16, j = 0..3
// Every input-channel count is a multiple of 16, so a lane's 8 consecutive
// k values are 8 consecutive channels of one tap: one 16-byte load.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scihpc {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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
// a BF16 value is > 0 (finite, or +inf) iff its bits read as a positive int16
__device__ __forceinline__ bool bf16_pos(bf16_t b) { return (int16_t)b > 0; }

__device__ __forceinline__ floatx4 mfma32(uint4 a, uint4 b, floatx4 c) {
  bf16x8 va, vb;
  __builtin_memcpy(&va, &a, 16);
  __builtin_memcpy(&vb, &b, 16);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, c, 0, 0, 0);
}

struct View {  // N(D)HWC channel slice
  bf16_t *p;
  int stride;  // channels of the underlying tensor
  int off;     // first channel of the slice
};

struct ConvShape {
  int n, d, h, w;  // images, depth (1 in 2-D), rows, columns
  int cin, cout;   // channels (multiples of 16)
};

template <int DIM>
struct Taps {
  static constexpr int T = DIM == 3 ? 27 : 9;
  // tap t -> offsets (dz, dy, dx) in {-1, 0, 1}; z-major, then y, then x
  static __device__ __forceinline__ void offsets(int t, int &dz, int &dy, int &dx) {
    const int r = DIM == 3 ? t % 9 : t;
    dz = DIM == 3 ? t / 9 - 1 : 0;
    dy = r / 3 - 1;
    dx = r % 3 - 1;
  }
};

constexpr int kConvThreads = 256;  // 4 wavefronts x 16 pixels = 64 pixels per workgroup

// y = act(conv(x, w) + b); w layout [cout][T][cin] (BF16), b FP32 (may be
// null).  mask: when given, x is read as x * (mask > 0) (ReLU derivative of
// the layer whose output gradient x is, for the backward-data use).  A
// workgroup computes NB output blocks of 16 channels, from block
// NB * blockIdx.y: a grid with gridDim.y > 1 splits the output channels of a
// small image over more workgroups (the coarse levels have few pixels and
// many channels).
template <int NB, int DIM, int G = 2>  // G: k-steps whose operands are loaded together
__global__ void __launch_bounds__(kConvThreads)
    k_conv(View x, View mask, const bf16_t *__restrict__ w, const float *__restrict__ b, View y, ConvShape s,
           int relu) {
  constexpr int T = Taps<DIM>::T;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int segs = (s.w + 63) / 64;
  int blk = blockIdx.x;
  const int seg = blk % segs;
  blk /= segs;
  const int row = blk % s.h;
  blk /= s.h;
  const int z = blk % s.d, img = blk / s.d;
  const int px = seg * 64 + wave * 16 + (lane & 15);  // pixel of this lane's A row
  const int kl = 8 * (lane >> 4);                    // k offset of this lane inside a step
  const int nb0 = NB * blockIdx.y;                   // first output block of this workgroup
  floatx4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int K = T * s.cin;
  // the operands of G consecutive k steps are loaded before their MFMAs, so
  // G loads are in flight at once instead of one dependent load per step
  for (int k0 = 0; k0 < K; k0 += 32 * G) {
    uint4 a[G], bw[G][NB];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int kk = k0 + 32 * g + kl;
      a[g] = uint4{0, 0, 0, 0};
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) bw[g][nb] = uint4{0, 0, 0, 0};
      if (kk < K) {
        const int t = kk / s.cin, c = kk - t * s.cin;
        int dz, dy, dx;
        Taps<DIM>::offsets(t, dz, dy, dx);
        const int zz = z + dz, yy = row + dy, xx = px + dx;
        if (zz >= 0 && zz < s.d && yy >= 0 && yy < s.h && xx >= 0 && xx < s.w) {
          const long pix = (((long)img * s.d + zz) * s.h + yy) * s.w + xx;
          a[g] = *reinterpret_cast<const uint4 *>(x.p + pix * x.stride + x.off + c);
          if (mask.p) {
            const uint4 m = *reinterpret_cast<const uint4 *>(mask.p + pix * mask.stride + mask.off + c);
            unsigned av[4] = {a[g].x, a[g].y, a[g].z, a[g].w};
            const unsigned mv[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (!bf16_pos((bf16_t)(mv[q] & 0xffffu))) av[q] &= 0xffff0000u;
              if (!bf16_pos((bf16_t)(mv[q] >> 16))) av[q] &= 0x0000ffffu;
            }
            a[g] = uint4{av[0], av[1], av[2], av[3]};
          }
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          bw[g][nb] =
              *reinterpret_cast<const uint4 *>(w + ((long)((nb0 + nb) * 16 + (lane & 15)) * T + t) * s.cin + c);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (k0 + 32 * g < K) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = mfma32(a[g], bw[g][nb], acc[nb]);
      }
    }
  }
  // epilogue: rows 4 (lane/16) + j of the wave's 16 pixels, column lane % 16
  const long rowpix = (((long)img * s.d + z) * s.h + row) * s.w;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = (nb0 + nb) * 16 + (lane & 15);
    const float bias = b ? b[co] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = seg * 64 + wave * 16 + 4 * (lane >> 4) + j;
      if (p >= s.w) continue;
      float v = acc[nb][j] + bias;
      if (relu) v = v > 0.f ? v : 0.f;
      y.p[(rowpix + p) * y.stride + y.off + co] = f_to_bf16(v);
    }
  }
}

// k_conv with the input tile staged in LDS: the workgroup's 64 output
// pixels need input rows (z, y) + {-1, 0, 1} (3 rows in 2-D, 9 in 3-D) of
// 66 pixels (one halo pixel each side) x cin channels; they are loaded
// once, coalesced (a row segment is contiguous in N(D)HWC), masked and
// zero-padded exactly as k_conv reads them, so every A fragment then comes
// from LDS instead of nine global reads of each input value.  LDS layout
// [row][pixel][channel], a pixel padded by 8 channels (16 bytes) to spread
// the banks; dynamic size kConvRows * 66 * (cin + 8) * 2 bytes.  The
// k-steps, MFMAs and epilogue are k_conv's, so the results are identical.
template <int DIM>
__host__ __device__ constexpr int conv_rows() {
  return DIM == 3 ? 9 : 3;
}
inline size_t conv_lds_bytes(int ndim, int cin) {
  return (size_t)(ndim == 3 ? 9 : 3) * 66 * (cin + 8) * sizeof(bf16_t);
}

// stage input rows (z, y) + {-1, 0, 1} x 66 pixels x cin of the workgroup's
// tile into LDS, masked and zero-padded as k_conv reads them
template <int DIM>
__device__ __forceinline__ void conv_stage_tile(View x, View mask, const ConvShape &s, int img, int z, int row, int seg,
                                                bf16_t *tile) {
  constexpr int R = conv_rows<DIM>();
  const int CP = s.cin + 8;  // LDS pixel pitch (channels)
  const int c8s = s.cin / 8;  // 16-byte chunks of a pixel
  for (int q = threadIdx.x; q < R * 66 * c8s; q += kConvThreads) {
    const int r = q / (66 * c8s), rem = q - r * 66 * c8s;
    const int p = rem / c8s, c = (rem - p * c8s) * 8;
    const int zz = z + (DIM == 3 ? r / 3 - 1 : 0), yy = row + (DIM == 3 ? r % 3 : r) - 1, xx = seg * 64 + p - 1;
    uint4 a = uint4{0, 0, 0, 0};
    if (zz >= 0 && zz < s.d && yy >= 0 && yy < s.h && xx >= 0 && xx < s.w) {
      const long pix = (((long)img * s.d + zz) * s.h + yy) * s.w + xx;
      a = *reinterpret_cast<const uint4 *>(x.p + pix * x.stride + x.off + c);
      if (mask.p) {
        const uint4 m = *reinterpret_cast<const uint4 *>(mask.p + pix * mask.stride + mask.off + c);
        unsigned av[4] = {a.x, a.y, a.z, a.w};
        const unsigned mv[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (!bf16_pos((bf16_t)(mv[u] & 0xffffu))) av[u] &= 0xffff0000u;
          if (!bf16_pos((bf16_t)(mv[u] >> 16))) av[u] &= 0x0000ffffu;
        }
        a = uint4{av[0], av[1], av[2], av[3]};
      }
    }
    *reinterpret_cast<uint4 *>(tile + (r * 66 + p) * CP + c) = a;
  }
}

template <int NB, int DIM, int G = 2>  // G: k-steps whose operands are loaded together
__global__ void __launch_bounds__(kConvThreads)
    k_conv_lds(View x, View mask, const bf16_t *__restrict__ w, const float *__restrict__ b, View y, ConvShape s,
               int relu) {
  constexpr int T = Taps<DIM>::T;
  extern __shared__ __align__(16) bf16_t tile[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int segs = (s.w + 63) / 64;
  int blk = blockIdx.x;
  const int seg = blk % segs;
  blk /= segs;
  const int row = blk % s.h;
  blk /= s.h;
  const int z = blk % s.d, img = blk / s.d;
  conv_stage_tile<DIM>(x, mask, s, img, z, row, seg, tile);
  const int CP = s.cin + 8;
  __syncthreads();
  const int pl = wave * 16 + (lane & 15);  // tile pixel of this lane's A row, before the tap's dx
  const int kl = 8 * (lane >> 4);
  const int nb0 = NB * blockIdx.y;
  floatx4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int K = T * s.cin;
  for (int k0 = 0; k0 < K; k0 += 32 * G) {
    uint4 a[G], bw[G][NB];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int kk = k0 + 32 * g + kl;
      a[g] = uint4{0, 0, 0, 0};
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) bw[g][nb] = uint4{0, 0, 0, 0};
      if (kk < K) {
        const int t = kk / s.cin, c = kk - t * s.cin;
        int dz, dy, dx;
        Taps<DIM>::offsets(t, dz, dy, dx);
        const int r = DIM == 3 ? (dz + 1) * 3 + dy + 1 : dy + 1;
        a[g] = *reinterpret_cast<const uint4 *>(tile + (r * 66 + pl + dx + 1) * CP + c);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          bw[g][nb] =
              *reinterpret_cast<const uint4 *>(w + ((long)((nb0 + nb) * 16 + (lane & 15)) * T + t) * s.cin + c);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (k0 + 32 * g < K) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = mfma32(a[g], bw[g][nb], acc[nb]);
      }
    }
  }
  const long rowpix = (((long)img * s.d + z) * s.h + row) * s.w;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = (nb0 + nb) * 16 + (lane & 15);
    const float bias = b ? b[co] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = seg * 64 + wave * 16 + 4 * (lane >> 4) + j;
      if (p >= s.w) continue;
      float v = acc[nb][j] + bias;
      if (relu) v = v > 0.f ? v : 0.f;
      y.p[(rowpix + p) * y.stride + y.off + co] = f_to_bf16(v);
    }
  }
}

// k_conv_lds for layers with several output blocks: the four wavefronts
// split the NB output blocks WS ways (WS = 2 or 4, NB a multiple of WS) and
// the tile's four 16-pixel groups 4 / WS ways.  Wavefront w computes the WS
// pixel groups from (w / WS) * WS on, for the NB / WS output blocks from
// (w % WS) * NB / WS on, so each weight fragment is loaded by 4 / WS
// wavefronts instead of all four.  Each (pixel group, output block) sees
// the same operands in the same k order as in k_conv: identical results.
template <int NB, int DIM, int WS, int G = 2>
__global__ void __launch_bounds__(kConvThreads)
    k_conv_ldsn(View x, View mask, const bf16_t *__restrict__ w, const float *__restrict__ b, View y, ConvShape s,
                int relu) {
  static_assert((WS == 2 || WS == 4) && NB % WS == 0, "the wavefronts split the output blocks evenly");
  constexpr int T = Taps<DIM>::T, NW = NB / WS, MP = WS;
  extern __shared__ __align__(16) bf16_t tile[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int segs = (s.w + 63) / 64;
  int blk = blockIdx.x;
  const int seg = blk % segs;
  blk /= segs;
  const int row = blk % s.h;
  blk /= s.h;
  const int z = blk % s.d, img = blk / s.d;
  conv_stage_tile<DIM>(x, mask, s, img, z, row, seg, tile);
  const int CP = s.cin + 8;
  __syncthreads();
  const int kl = 8 * (lane >> 4);
  const int nbw = NB * blockIdx.y + NW * (wave % WS);  // first output block of this wavefront
  const int pg0 = MP * (wave / WS);                      // first 16-pixel group of this wavefront
  floatx4 acc[MP][NW];
#pragma unroll
  for (int m = 0; m < MP; ++m)
#pragma unroll
    for (int nb = 0; nb < NW; ++nb) acc[m][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int K = T * s.cin;
  for (int k0 = 0; k0 < K; k0 += 32 * G) {
    uint4 a[G][MP], bw[G][NW];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int kk = k0 + 32 * g + kl;
#pragma unroll
      for (int m = 0; m < MP; ++m) a[g][m] = uint4{0, 0, 0, 0};
#pragma unroll
      for (int nb = 0; nb < NW; ++nb) bw[g][nb] = uint4{0, 0, 0, 0};
      if (kk < K) {
        const int t = kk / s.cin, c = kk - t * s.cin;
        int dz, dy, dx;
        Taps<DIM>::offsets(t, dz, dy, dx);
        const int r = DIM == 3 ? (dz + 1) * 3 + dy + 1 : dy + 1;
#pragma unroll
        for (int m = 0; m < MP; ++m)
          a[g][m] = *reinterpret_cast<const uint4 *>(tile + (r * 66 + (pg0 + m) * 16 + (lane & 15) + dx + 1) * CP + c);
#pragma unroll
        for (int nb = 0; nb < NW; ++nb)
          bw[g][nb] =
              *reinterpret_cast<const uint4 *>(w + ((long)((nbw + nb) * 16 + (lane & 15)) * T + t) * s.cin + c);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (k0 + 32 * g < K) {
#pragma unroll
        for (int m = 0; m < MP; ++m)
#pragma unroll
          for (int nb = 0; nb < NW; ++nb) acc[m][nb] = mfma32(a[g][m], bw[g][nb], acc[m][nb]);
      }
    }
  }
  const long rowpix = (((long)img * s.d + z) * s.h + row) * s.w;
#pragma unroll
  for (int nb = 0; nb < NW; ++nb) {
    const int co = (nbw + nb) * 16 + (lane & 15);
    const float bias = b ? b[co] : 0.f;
#pragma unroll
    for (int m = 0; m < MP; ++m)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = seg * 64 + (pg0 + m) * 16 + 4 * (lane >> 4) + j;
        if (p >= s.w) continue;
        float v = acc[m][nb][j] + bias;
        if (relu) v = v > 0.f ? v : 0.f;
        y.p[(rowpix + p) * y.stride + y.off + co] = f_to_bf16(v);
      }
  }
}

// k_conv for the large, narrow layers (full-resolution images, few
// channels): a workgroup loads the layer's weight fragments for all KS
// k-steps into registers once, then computes kConvTilesPerWG pixel tiles,
// each with all of its KS input fragments loaded before the first MFMA, so
// a tile waits on one round of loads instead of one per two k-steps.  The
// k-steps run in k_conv's order with the same operands: identical results.
constexpr int kConvTilesPerWG = 4;
// the deepest reductions (KS >= 14 k-steps: the 48 -> 16 decoder layer) take
// 8 tiles a workgroup (55.0 -> 50.1 us at 256x512, batch 4); 8 for every
// layer cost the 128x256 layers, whose 2048 tiles then fill 256 workgroups
template <int KS>
constexpr int rw_tiles() {
  return KS >= 14 ? 2 * kConvTilesPerWG : kConvTilesPerWG;
}

template <int NB, int DIM, int KS>
__global__ void __launch_bounds__(kConvThreads)
    k_conv_rw(View x, View mask, const bf16_t *__restrict__ w, const float *__restrict__ b, View y, ConvShape s,
              int relu, int ntiles) {
  constexpr int T = Taps<DIM>::T;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kl = 8 * (lane >> 4);
  const int K = T * s.cin;
  uint4 bw[KS][NB];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int kk = 32 * ks + kl;
    const int t = kk / s.cin, c = kk - t * s.cin;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      bw[ks][nb] = kk < K ? *reinterpret_cast<const uint4 *>(w + ((long)(nb * 16 + (lane & 15)) * T + t) * s.cin + c)
                          : uint4{0, 0, 0, 0};
  }
  float bias[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) bias[nb] = b ? b[nb * 16 + (lane & 15)] : 0.f;
  const int segs = (s.w + 63) / 64;
  const int t0 = blockIdx.x * rw_tiles<KS>(), t1 = min(t0 + rw_tiles<KS>(), ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    int blk = tile;
    const int seg = blk % segs;
    blk /= segs;
    const int row = blk % s.h;
    blk /= s.h;
    const int z = blk % s.d, img = blk / s.d;
    const int px = seg * 64 + wave * 16 + (lane & 15);
    uint4 a[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kk = 32 * ks + kl;
      a[ks] = uint4{0, 0, 0, 0};
      if (kk < K) {
        const int t = kk / s.cin, c = kk - t * s.cin;
        int dz, dy, dx;
        Taps<DIM>::offsets(t, dz, dy, dx);
        const int zz = z + dz, yy = row + dy, xx = px + dx;
        if (zz >= 0 && zz < s.d && yy >= 0 && yy < s.h && xx >= 0 && xx < s.w) {
          const long pix = (((long)img * s.d + zz) * s.h + yy) * s.w + xx;
          a[ks] = *reinterpret_cast<const uint4 *>(x.p + pix * x.stride + x.off + c);
          if (mask.p) {
            const uint4 m = *reinterpret_cast<const uint4 *>(mask.p + pix * mask.stride + mask.off + c);
            unsigned av[4] = {a[ks].x, a[ks].y, a[ks].z, a[ks].w};
            const unsigned mv[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (!bf16_pos((bf16_t)(mv[q] & 0xffffu))) av[q] &= 0xffff0000u;
              if (!bf16_pos((bf16_t)(mv[q] >> 16))) av[q] &= 0x0000ffffu;
            }
            a[ks] = uint4{av[0], av[1], av[2], av[3]};
          }
        }
      }
    }
    floatx4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[nb] = mfma32(a[ks], bw[ks][nb], acc[nb]);
    const long rowpix = (((long)img * s.d + z) * s.h + row) * s.w;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = nb * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = seg * 64 + wave * 16 + 4 * (lane >> 4) + j;
        if (p >= s.w) continue;
        float v = acc[nb][j] + bias[nb];
        if (relu) v = v > 0.f ? v : 0.f;
        y.p[(rowpix + p) * y.stride + y.off + co] = f_to_bf16(v);
      }
    }
  }
}

// Weight and bias gradient of y = relu(conv(x, w) + b):
//   dW[co][t][ci] = sum_p gz[p][co] x[p + off(t)][ci],  db[co] = sum_p gz[p][co]
// with gz = g * (mask > 0) (mask = the layer's stored output).  A = gz^T
// (output channels x pixels), B = x shifted by the tap (pixels x input
// channels), K = pixels in steps of 32.  A workgroup owns one 16 x 16
// channel block, one group of 9 taps (the 3-D kernel has three, one per
// z offset) and one pixel chunk; each wavefront takes 32 consecutive pixels
// of every 128.  The tiles reach the matrix cores through LDS: a lane loads
// 8 channels of one pixel (16 bytes) and stores them transposed, so an
// operand read is 8 pixels of one channel (one 16-byte LDS read).  The four
// wavefronts are summed through LDS and the workgroup writes its partial
// (no atomics) to
//   part[chunk][co][t][ci], part[chunk][cout*T*cin + co]   (bias, from the
// workgroups of cc == 0 and tap group 0 only)
// which k_sum_rows_2d reduces over chunks in a fixed order.
constexpr int kWgPad = 40;

// LDS exchange among the lanes of one wavefront: a wavefront's LDS
// operations complete in order, so ordering the compiler's accesses (the
// wavefront-scope fences) and keeping the lanes together is enough; no
// workgroup barrier
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}  // LDS row length (BF16) of a 32-pixel tile row; 80 bytes keep 16-byte reads aligned

template <int DIM>
__global__ void __launch_bounds__(kConvThreads)
    k_conv_wgrad(View x, View g, View mask, float *__restrict__ part, ConvShape s, int chunk_pix) {
  constexpr int T = Taps<DIM>::T;
  constexpr int NG = T / 9;  // tap groups
  __shared__ __align__(16) bf16_t tg[4][16][kWgPad];
  __shared__ __align__(16) bf16_t tx[4][16][kWgPad];
  __shared__ float red[4][36][64];
  __shared__ float bred[4][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cbs = s.cout / 16, ccs = s.cin / 16;
  int blk = blockIdx.x;
  const int cb = blk % cbs;
  blk /= cbs;
  const int cc = blk % ccs;
  blk /= ccs;
  const int grp = blk % NG;
  const int chunk = blk / NG;
  const long npix = (long)s.n * s.d * s.h * s.w;
  const long p_begin = (long)chunk * chunk_pix;
  const long p_end = p_begin + chunk_pix < npix ? p_begin + chunk_pix : npix;
  floatx4 acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int sp = lane >> 1, half = 8 * (lane & 1);  // staging: pixel of the tile, channel half
  const int kl = 8 * (lane >> 4);                   // operand read: first pixel of this lane
  // the tiles tg[wave] and tx[wave] are each wavefront's own: the loop
  // synchronises within a wavefront only.  The loads of the next 128 pixels
  // (every tap's shifted input, the output gradient and its mask) are issued
  // before the current pixels are transposed and multiplied, so their
  // latency overlaps that work.
  auto load = [&](long base, uint4(&xv)[9], uint4 &gv, uint4 &mv) {
    const long p = base + wave * 32 + sp;
    const bool pv = p < p_end;
    int px = 0, py = 0, pz = 0;
    if (pv) {
      px = (int)(p % s.w);
      const long r = p / s.w;
      py = (int)(r % s.h);
      pz = (int)((r / s.h) % s.d);
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      int dz, dy, dx;
      Taps<DIM>::offsets(grp * 9 + t, dz, dy, dx);
      xv[t] = uint4{0, 0, 0, 0};
      const int zz = pz + dz, yy = py + dy, xx = px + dx;
      if (pv && zz >= 0 && zz < s.d && yy >= 0 && yy < s.h && xx >= 0 && xx < s.w) {
        const long q = p + ((long)dz * s.h + dy) * s.w + dx;
        xv[t] = *reinterpret_cast<const uint4 *>(x.p + q * x.stride + x.off + cc * 16 + half);
      }
    }
    gv = uint4{0, 0, 0, 0};
    mv = uint4{0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u};  // BF16 1.0: nothing masked
    if (pv) {
      gv = *reinterpret_cast<const uint4 *>(g.p + p * g.stride + g.off + cb * 16 + half);
      if (mask.p) mv = *reinterpret_cast<const uint4 *>(mask.p + p * mask.stride + mask.off + cb * 16 + half);
    }
  };
  uint4 xv[9], gv, mv;
  load(p_begin, xv, gv, mv);
  for (long base = p_begin; base < p_end; base += 128) {
    uint4 xn[9], gn, mn;
    if (base + 128 < p_end) {
      load(base + 128, xn, gn, mn);
    } else {
#pragma unroll
      for (int t = 0; t < 9; ++t) xn[t] = uint4{0, 0, 0, 0};
      gn = mn = uint4{0, 0, 0, 0};
    }
    {  // gz tile [co][pixel], the ReLU mask applied
      unsigned gw[4] = {gv.x, gv.y, gv.z, gv.w};
      if (mask.p) {
        const unsigned mw[4] = {mv.x, mv.y, mv.z, mv.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!bf16_pos((bf16_t)(mw[q] & 0xffffu))) gw[q] &= 0xffff0000u;
          if (!bf16_pos((bf16_t)(mw[q] >> 16))) gw[q] &= 0x0000ffffu;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        tg[wave][half + 2 * q][sp] = (bf16_t)(gw[q] & 0xffffu);
        tg[wave][half + 2 * q + 1][sp] = (bf16_t)(gw[q] >> 16);
      }
    }
    wave_lds_sync();  // tg is this wavefront's own
    const uint4 a = *reinterpret_cast<const uint4 *>(&tg[wave][lane & 15][kl]);
    {  // bias: this lane's 8 pixels of channel lane % 16
      const unsigned av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) bsum += bf16_to_f((bf16_t)(av[q] & 0xffffu)) + bf16_to_f((bf16_t)(av[q] >> 16));
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const unsigned xw[4] = {xv[t].x, xv[t].y, xv[t].z, xv[t].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        tx[wave][half + 2 * q][sp] = (bf16_t)(xw[q] & 0xffffu);
        tx[wave][half + 2 * q + 1][sp] = (bf16_t)(xw[q] >> 16);
      }
      wave_lds_sync();
      const uint4 bb = *reinterpret_cast<const uint4 *>(&tx[wave][lane & 15][kl]);
      acc[t] = mfma32(a, bb, acc[t]);
      wave_lds_sync();  // the next tap overwrites tx
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) xv[t] = xn[t];
    gv = gn;
    mv = mn;
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
  const long rowlen = (long)s.cout * T * s.cin + s.cout;
  float *dst = part + (long)chunk * rowlen;
  for (int e = threadIdx.x; e < 36 * 64; e += kConvThreads) {
    const int tj = e >> 6, l = e & 63;
    const int t = tj >> 2, j = tj & 3;
    const float v = red[0][tj][l] + red[1][tj][l] + red[2][tj][l] + red[3][tj][l];
    const int o = cb * 16 + 4 * (l >> 4) + j, c = cc * 16 + (l & 15);
    dst[((long)o * T + grp * 9 + t) * s.cin + c] = v;
  }
  if (cc == 0 && grp == 0 && threadIdx.x < 16)
    dst[(long)s.cout * T * s.cin + cb * 16 + threadIdx.x] =
        bred[0][threadIdx.x] + bred[1][threadIdx.x] + bred[2][threadIdx.x] + bred[3][threadIdx.x];
}

}  // namespace scihpc
