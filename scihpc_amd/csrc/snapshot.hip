// snapshot.hip -- legacy-VTK snapshot writer behind the C ABI of
// include/scihpc_hip.h, byte-for-byte the files of vtkWriter
// (vtkWriter.cpp:18-111): ASCII header, BINARY STRUCTURED_POINTS,
// CELL_DATA, FP64 big-endian values in k-j-i order (x fastest).
//
// The GPU does the reformatting: one LDS-tiled transpose kernel gathers the
// interior of a padded field (data[i][j][k], k fastest) into x-fastest
// order and byte-swaps every value, writing the file image of the field
// into a device staging buffer.  The image goes to pinned host memory in
// chunks and a writer thread appends each chunk to the file once its copy
// has landed, so the next chunk's copy overlaps the previous chunk's write
// and the caller's stream is never blocked on the disk.
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>

#include "common.hpp"

namespace scihpc {
namespace {

constexpr int kVtkGhc = 3;
constexpr int kTile = 32;                  // transpose tile edge (doubles)
constexpr int kTileRows = 8;               // threads along the tile's rows
constexpr size_t kChunk = size_t(16) << 20;  // bytes per pinned slot
constexpr int kSlots = 3;

__device__ __forceinline__ double be64(double x) {
  const unsigned long long b = __builtin_bswap64(__double_as_longlong(x));
  return __longlong_as_double((long long)b);
}

// Layout of one padded field: interior sizes, padded strides, ghost offsets,
// the contiguous input axis `ca` (the last real axis) and the remaining
// batch axis `ba` (y in 3-D; the unit z axis in 1-D and 2-D).
struct VtkShape {
  int n[3];
  long s[3];
  int off[3];
  int ca;   // contiguous input axis (1 in 2-D, 2 in 3-D)
  int ba;   // batch axis (2 in 2-D with extent 1, 1 in 3-D)
};

// Transposes NC components of a padded field into the VTK image:
// out[NC * (i + nx*(j + ny*k)) + c] = bswap(in_c[(i+o0)*s0 + (j+o1)*s1 + (k+o2)*s2]).
// Grid: x tiles along i, y tiles along the contiguous axis, z = batch index.
template <int NC>
__global__ void __launch_bounds__(kTile *kTileRows)
    k_vtk_pack(const double *__restrict__ f0, const double *__restrict__ f1,
               const double *__restrict__ f2, double *__restrict__ out, VtkShape sh) {
  __shared__ double tile[NC][kTile][kTile + 1];
  const double *fs[3] = {f0, f1, f2};
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int i0 = blockIdx.x * kTile, c0 = blockIdx.y * kTile, b = blockIdx.z;
  const int nx = sh.n[0], nc = sh.n[sh.ca];
  int idx[3];
  idx[sh.ba] = b;
  // load: rows are i, columns the contiguous axis (coalesced reads)
#pragma unroll
  for (int r = 0; r < kTile / kTileRows; ++r) {
    const int i = i0 + ty + r * kTileRows, c = c0 + tx;
    if (i < nx && c < nc) {
      idx[0] = i;
      idx[sh.ca] = c;
      const long p = (long)(idx[0] + sh.off[0]) * sh.s[0] + (long)(idx[1] + sh.off[1]) * sh.s[1] +
                     (long)(idx[2] + sh.off[2]) * sh.s[2];
#pragma unroll
      for (int q = 0; q < NC; ++q) tile[q][ty + r * kTileRows][tx] = fs[q][p];
    }
  }
  __syncthreads();
  // store: x fastest (coalesced writes)
#pragma unroll
  for (int r = 0; r < kTile / kTileRows; ++r) {
    const int c = c0 + ty + r * kTileRows, i = i0 + tx;
    if (i < nx && c < nc) {
      idx[0] = i;
      idx[sh.ca] = c;
      const long o = (long)idx[0] + (long)sh.n[0] * (idx[1] + (long)sh.n[1] * idx[2]);
#pragma unroll
      for (int q = 0; q < NC; ++q) out[NC * o + q] = be64(tile[q][tx][ty + r * kTileRows]);
    }
  }
}

// std::ostream's default formatting of a double (precision 6, %g), which
// is what vtkWriter::create streams into the header.
std::string g6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%g", v);
  return buf;
}

}  // namespace
}  // namespace scihpc

using namespace scihpc;

struct scihpc_vtk {
  int ndim = 0;
  VtkShape sh{};
  long cells = 0;   // interior cells
  long padded = 0;  // padded cells of one field
  double origin[3]{}, spacing[3]{};
  int vtk_dim[3]{};
  std::string prefix;
  hipStream_t stream = nullptr;
  hipEvent_t packed = nullptr;  // the last pack kernel has read its fields
  double *img = nullptr;      // device file image of one field (NC * cells)
  double *dstage = nullptr;   // device copy of host-resident padded fields
  long dstage_fields = 0;
  FILE *file = nullptr;
  std::string io_error;

  struct Slot {
    unsigned char *host = nullptr;
    hipEvent_t ready = nullptr;
    bool busy = false;
  };
  Slot slot[kSlots];
  int next_slot = 0;

  // writer thread: appends text and landed chunks to `file` in order
  struct Job {
    int kind;  // 0 text, 1 slot chunk, 2 close
    std::string text;
    int slot;
    size_t bytes;
  };
  std::thread writer;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  bool quit = false;
  long pending = 0;  // jobs queued or running

  void run() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return quit || !jobs.empty(); });
        if (jobs.empty()) return;
        j = std::move(jobs.front());
        jobs.pop_front();
      }
      FILE *f;
      {
        std::lock_guard<std::mutex> lk(mu);
        f = file;
      }
      std::string err;
      if (j.kind == 0) {
        if (f && std::fwrite(j.text.data(), 1, j.text.size(), f) != j.text.size()) err = "short write";
      } else if (j.kind == 1) {
        const hipError_t e = hipEventSynchronize(slot[j.slot].ready);
        if (e != hipSuccess) err = std::string("chunk copy: ") + hipGetErrorString(e);
        else if (f && std::fwrite(slot[j.slot].host, 1, j.bytes, f) != j.bytes) err = "short write";
      } else {
        if (f && std::fclose(f) != 0) err = "close failed";
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!err.empty() && io_error.empty()) io_error = err;
        if (j.kind == 1) slot[j.slot].busy = false;
        if (j.kind == 2) file = nullptr;
        --pending;
      }
      cv.notify_all();
    }
  }

  void push(Job j) {
    {
      std::lock_guard<std::mutex> lk(mu);
      jobs.push_back(std::move(j));
      ++pending;
    }
    cv.notify_all();
  }

  void text(const std::string &s) { push(Job{0, s, 0, 0}); }

  // Waits for every queued job; returns and clears the first I/O error.
  std::string drain() {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return pending == 0; });
    std::string e = io_error;
    io_error.clear();
    return e;
  }

  // Packs NC padded device fields into `img`, then streams the image to the
  // file through the pinned slots.
  template <int NC>
  void emit(const double *f0, const double *f1, const double *f2) {
    dim3 grid(ceil_div(sh.n[0], kTile), ceil_div(sh.n[sh.ca], kTile), sh.n[sh.ba]);
    hipLaunchKernelGGL(k_vtk_pack<NC>, grid, dim3(kTile * kTileRows), 0, stream, f0, f1, f2, img, sh);
    SCIHPC_CHECK_LAUNCH();
    // the caller may reuse its fields as soon as the call returns
    SCIHPC_HIP(hipEventRecord(packed, stream));
    SCIHPC_HIP(hipEventSynchronize(packed));
    const size_t total = sizeof(double) * NC * (size_t)cells;
    const unsigned char *src = reinterpret_cast<const unsigned char *>(img);
    for (size_t done = 0; done < total; done += kChunk) {
      const size_t bytes = std::min(kChunk, total - done);
      const int s = next_slot;
      next_slot = (next_slot + 1) % kSlots;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !slot[s].busy; });
        slot[s].busy = true;
      }
      SCIHPC_HIP(hipMemcpyAsync(slot[s].host, src + done, bytes, hipMemcpyDeviceToHost, stream));
      SCIHPC_HIP(hipEventRecord(slot[s].ready, stream));
      push(Job{1, std::string(), s, bytes});
    }
  }

  const double *device_field(const double *p, int on_device, int q) {
    if (on_device) return p;
    SCIHPC_HIP(hipMemcpyAsync(dstage + q * padded, p, sizeof(double) * padded, hipMemcpyHostToDevice, stream));
    return dstage + q * padded;
  }

  void ensure_stage(int nfields) {
    if (dstage_fields >= nfields) return;
    SCIHPC_HIP(hipStreamSynchronize(stream));
    if (dstage) SCIHPC_HIP(hipFree(dstage));
    dstage = nullptr;
    SCIHPC_HIP(hipMalloc(&dstage, sizeof(double) * padded * nfields));
    SCIHPC_HIP(hipMemsetAsync(dstage, 0, sizeof(double) * padded * nfields, stream));
    dstage_fields = nfields;
  }
};

extern "C" int scihpc_vtk_destroy(scihpc_vtk *h);

extern "C" int scihpc_vtk_create(int ndim, int nx, int ny, int nz, const double *origin,
                                 const double *spacing, const char *prefix, scihpc_vtk **out) {
  return guarded([&] {
    SCIHPC_REQUIRE(out, "out is null");
    *out = nullptr;
    SCIHPC_REQUIRE(ndim >= 1 && ndim <= 3, "ndim must be 1, 2 or 3");
    SCIHPC_REQUIRE(origin && spacing && prefix, "origin, spacing and prefix are required");
    const int n[3] = {nx, ndim > 1 ? ny : 1, ndim > 2 ? nz : 1};
    for (int a = 0; a < ndim; ++a) SCIHPC_REQUIRE(n[a] > 0, "cell counts must be positive");
    auto *h = new scihpc_vtk;
    try {
      h->ndim = ndim;
      h->prefix = prefix;
      int N[3];
      for (int a = 0; a < 3; ++a) {
        h->sh.n[a] = n[a];
        N[a] = a < ndim ? n[a] + 2 * kVtkGhc : 1;
        h->sh.off[a] = a < ndim ? kVtkGhc : 0;
        // structured_grid: vtk_info.n = cells + 1 on a real axis, 1 otherwise;
        // dy = dz = 0 below ndim (structured_grid.cpp:7-41, :43-84)
        h->vtk_dim[a] = a < ndim ? n[a] + 1 : 1;
        h->origin[a] = a < ndim ? origin[a] : 0.0;
        h->spacing[a] = a < ndim ? spacing[a] : 0.0;
      }
      h->sh.s[2] = 1;
      h->sh.s[1] = N[2];
      h->sh.s[0] = (long)N[1] * N[2];
      h->sh.ca = ndim == 3 ? 2 : 1;
      h->sh.ba = ndim == 3 ? 1 : 2;
      h->cells = (long)n[0] * n[1] * n[2];
      h->padded = (long)N[0] * N[1] * N[2];
      SCIHPC_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
      SCIHPC_HIP(hipEventCreateWithFlags(&h->packed, hipEventDisableTiming));
      SCIHPC_HIP(hipMalloc(&h->img, sizeof(double) * 3 * h->cells));
      for (auto &s : h->slot) {
        SCIHPC_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.host), kChunk, hipHostMallocDefault));
        SCIHPC_HIP(hipEventCreateWithFlags(&s.ready, hipEventDisableTiming));
      }
      h->writer = std::thread([h] { h->run(); });
    } catch (...) {
      scihpc_vtk_destroy(h);
      throw;
    }
    *out = h;
  });
}

extern "C" int scihpc_vtk_destroy(scihpc_vtk *h) {
  if (!h) return SCIHPC_OK;
  if (h->writer.joinable()) {
    if (h->file) h->push(scihpc_vtk::Job{2, std::string(), 0, 0});
    {
      std::lock_guard<std::mutex> lk(h->mu);
      h->quit = true;
    }
    h->cv.notify_all();
    h->writer.join();
  }
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &s : h->slot) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.ready) (void)hipEventDestroy(s.ready);
  }
  if (h->packed) (void)hipEventDestroy(h->packed);
  if (h->img) (void)hipFree(h->img);
  if (h->dstage) (void)hipFree(h->dstage);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SCIHPC_OK;
}

extern "C" int scihpc_vtk_begin(scihpc_vtk *h, unsigned int id) {
  return guarded([&] {
    SCIHPC_REQUIRE(h, "null handle");
    const std::string err = h->drain();
    if (h->file) {
      h->push(scihpc_vtk::Job{2, std::string(), 0, 0});
      (void)h->drain();
    }
    if (!err.empty()) throw Error(SCIHPC_IO_ERROR, "vtk: " + err);
    const std::string path = h->prefix + std::to_string(id) + ".vtk";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(SCIHPC_IO_ERROR, "vtk: cannot open " + path);
    {
      std::lock_guard<std::mutex> lk(h->mu);
      h->file = f;
    }
    // vtkWriter::create (vtkWriter.cpp:18-36)
    std::string t = "# vtk DataFile Version 3.0\nvtk output\nBINARY\nDATASET STRUCTURED_POINTS\n";
    t += "DIMENSIONS " + std::to_string(h->vtk_dim[0]) + " " + std::to_string(h->vtk_dim[1]) + " " +
         std::to_string(h->vtk_dim[2]) + "\n";
    t += "ORIGIN " + g6(h->origin[0]) + " " + g6(h->origin[1]) + " " + g6(h->origin[2]) + "\n";
    t += "SPACING " + g6(h->spacing[0]) + " " + g6(h->spacing[1]) + " " + g6(h->spacing[2]) + "\n";
    long size = 1;
    for (int a = 0; a < 3; ++a)
      if (h->vtk_dim[a] > 1) size *= h->vtk_dim[a] - 1;
    t += "CELL_DATA " + std::to_string(size) + "\n";
    h->text(t);
  });
}

extern "C" int scihpc_vtk_add_scalar(scihpc_vtk *h, const char *name, const double *field,
                                     int on_device) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && name && field, "null argument");
    SCIHPC_REQUIRE(h->file, "vtk: no open file (call scihpc_vtk_begin first)");
    if (!on_device) h->ensure_stage(1);
    const double *d = h->device_field(field, on_device, 0);
    // vtkWriter::add_scalar (vtkWriter.cpp:38-52)
    h->text(std::string("SCALARS ") + name + " double 1\nLOOKUP_TABLE default\n");
    h->emit<1>(d, d, d);
  });
}

extern "C" int scihpc_vtk_add_vector(scihpc_vtk *h, const char *name, const double *u,
                                     const double *v, const double *w, int on_device) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && name && u && v && w, "null argument");
    SCIHPC_REQUIRE(h->file, "vtk: no open file (call scihpc_vtk_begin first)");
    if (!on_device) h->ensure_stage(3);
    const double *du = h->device_field(u, on_device, 0);
    const double *dv = h->device_field(v, on_device, 1);
    const double *dw = h->device_field(w, on_device, 2);
    // vtkWriter::add_vector (vtkWriter.cpp:54-76)
    h->text(std::string("VECTORS ") + name + " double\n");
    h->emit<3>(du, dv, dw);
  });
}

extern "C" int scihpc_vtk_close(scihpc_vtk *h) {
  return guarded([&] {
    SCIHPC_REQUIRE(h, "null handle");
    if (h->file) h->push(scihpc_vtk::Job{2, std::string(), 0, 0});
    const std::string err = h->drain();
    if (!err.empty()) throw Error(SCIHPC_IO_ERROR, "vtk: " + err);
  });
}

extern "C" int scihpc_vtk_time_pack(scihpc_vtk *h, int ncomp, int iters, double *ms) {
  return guarded([&] {
    SCIHPC_REQUIRE(h && ms && iters > 0, "bad argument");
    SCIHPC_REQUIRE(ncomp == 1 || ncomp == 3, "ncomp must be 1 or 3");
    h->ensure_stage(3);
    const double *u = h->dstage, *v = h->dstage + h->padded, *w = h->dstage + 2 * h->padded;
    hipEvent_t a, b;
    SCIHPC_HIP(hipEventCreate(&a));
    SCIHPC_HIP(hipEventCreate(&b));
    dim3 grid(ceil_div(h->sh.n[0], kTile), ceil_div(h->sh.n[h->sh.ca], kTile), h->sh.n[h->sh.ba]);
    auto launch = [&] {
      if (ncomp == 1)
        hipLaunchKernelGGL(k_vtk_pack<1>, grid, dim3(kTile * kTileRows), 0, h->stream, u, u, u, h->img, h->sh);
      else
        hipLaunchKernelGGL(k_vtk_pack<3>, grid, dim3(kTile * kTileRows), 0, h->stream, u, v, w, h->img, h->sh);
    };
    launch();
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipEventRecord(a, h->stream));
    for (int it = 0; it < iters; ++it) launch();
    SCIHPC_CHECK_LAUNCH();
    SCIHPC_HIP(hipEventRecord(b, h->stream));
    SCIHPC_HIP(hipEventSynchronize(b));
    float t = 0;
    SCIHPC_HIP(hipEventElapsedTime(&t, a, b));
    *ms = t / iters;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  });
}
