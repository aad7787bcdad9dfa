// common.hpp -- shared HIP plumbing for the gfx950 kernels: status codes,
// error capture for the C ABI, wavefront/block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>

#include "scihpc_hip.h"

namespace scihpc {

// Thrown inside the library, converted to a status code at the C ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define SCIHPC_HIP(call)                                                      \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::scihpc::Error(SCIHPC_HIP_ERROR,                                 \
                            std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define SCIHPC_CHECK_LAUNCH() SCIHPC_HIP(hipGetLastError())

#define SCIHPC_REQUIRE(cond, msg)                                             \
  do {                                                                        \
    if (!(cond)) throw ::scihpc::Error(SCIHPC_BAD_ARG, msg);                  \
  } while (0)

// Runs `body` and maps exceptions to status codes; used by every extern "C"
// entry point so that nothing throws across the ABI.
template <typename F>
int guarded(F &&body) {
  try {
    body();
    return SCIHPC_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return SCIHPC_HIP_ERROR;
  } catch (...) {
    set_last_error("unknown exception");
    return SCIHPC_HIP_ERROR;
  }
}

constexpr int kWave = 64;  // CDNA wavefront

// Sum across the 64 lanes of a wavefront (butterfly over xor offsets; every
// lane ends with the total).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum for blockDim.x <= 1024; result valid in thread 0.  The
// summation order is fixed by the launch shape, so results are
// deterministic run to run.
template <typename T>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T part[1024 / kWave];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  v = wave_sum(v);
  if (lane == 0) part[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + kWave - 1) / kWave;
    for (int w = 0; w < nw; ++w) s += part[w];
  }
  __syncthreads();
  return s;
}

// block_sum with the total returned to every thread
template <typename T>
__device__ __forceinline__ T block_allsum(T v) {
  __shared__ T tot;
  const T s = block_sum(v);
  if (threadIdx.x == 0) tot = s;
  __syncthreads();
  const T r = tot;
  __syncthreads();  // tot is reused by the next call
  return r;
}

// two block_allsum in one pass (one LDS exchange instead of two; the same
// summation order as block_sum for each)
template <typename T>
__device__ __forceinline__ void block_allsum2(T &a, T &b) {
  __shared__ T pa[1024 / kWave], pb[1024 / kWave];
  __shared__ T ta, tb;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane == 0) {
    pa[wid] = a;
    pb[wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + kWave - 1) / kWave;
    T sa = 0, sb = 0;
    for (int w = 0; w < nw; ++w) {
      sa += pa[w];
      sb += pb[w];
    }
    ta = sa;
    tb = sb;
  }
  __syncthreads();
  a = ta;
  b = tb;
  __syncthreads();  // ta, tb are reused by the next call
}

// block_sum of N values in one LDS exchange (each summed in block_sum's
// order); results valid in thread 0
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N]) {
  __shared__ double part[N][1024 / kWave];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) part[q][wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double s = 0;
      for (int w = 0; w < nw; ++w) s += part[q][w];
      v[q] = s;
    }
  }
  __syncthreads();
}

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace scihpc
