// learned.hpp -- how the multigrid solve reaches the learned pressure guess
// (U-Net, unet.hip).  The solver owns the features, the prolongation of the
// guess, the acceptance test on the FP64 residual and the fallback to plain
// multigrid; the network only maps features to a coarse-level guess.
#pragma once

#include <hip/hip_runtime.h>

struct scihpc_unet;

namespace scihpc {

// out[pixels] = network(feat[pixels][2]) in normalised units, stream ordered
void unet_infer(scihpc_unet *net, const float *feat, float *out, hipStream_t stream);
int unet_ndim(const scihpc_unet *net);
int unet_depth(const scihpc_unet *net);
int unet_height(const scihpc_unet *net);
int unet_width(const scihpc_unet *net);
float unet_out_scale(const scihpc_unet *net);

struct LearnedHook {
  scihpc_unet *net = nullptr;
  int level = 1;  // multigrid level whose grid the network works on
  bool active() const { return net != nullptr; }
};

}  // namespace scihpc
