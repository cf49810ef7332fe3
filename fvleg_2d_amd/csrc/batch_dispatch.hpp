// Launch table of the batched persistent stage kernel (per precision x tile).
#pragma once

#include "batch_kernel.hpp"
#include "stage_dispatch.hpp"

namespace cato {

using BatchLauncher = void (*)(dim3 grid, hipStream_t stream, const BatchStageArgs &args);
BatchLauncher batch_launcher_p0(int tile);
BatchLauncher batch_launcher_p1(int tile);
inline BatchLauncher batch_launcher(int prec, int tile) {
  return prec == CATO_PRECISION_FP32_FLUX ? batch_launcher_p1(tile) : batch_launcher_p0(tile);
}

}  // namespace cato
