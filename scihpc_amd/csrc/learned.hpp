// learned.hpp -- hook through which the multigrid solve asks the learned
// preconditioner (U-Net, unet.hip) for a pressure guess.  The solver owns
// the acceptance test (FP64 residual) and the fallback to plain multigrid.
#pragma once

#include <hip/hip_runtime.h>

namespace scihpc {

struct LearnedModel;  // defined in unet.hip

struct LearnedHook {
  LearnedModel *model = nullptr;
  const double *rho = nullptr;  // level-0 density (MG layout, device)
  bool active() const { return model != nullptr; }
  // x = network(b, log rho), FP64 in/out on the device, stream ordered
  void apply(const double *b, double *x, hipStream_t stream);
};

}  // namespace scihpc
