// Batched persistent stage kernel instances, precision id 1.
#include "batch_dispatch.hpp"

namespace cato {
namespace {
using R1 = f32x;  // FP32-flux form arithmetic
template <int TX, int TY> void launch(dim3 g, hipStream_t s, const BatchStageArgs &a) {
  hipLaunchKernelGGL((batch_stage_kernel<R1, TX, TY>), g, dim3(kThreads), 0, s, a);
}
}  // namespace

BatchLauncher batch_launcher_p1(int tile) {
  switch (tile) {
    case 0: return &launch<kTiles[0].tx, kTiles[0].ty>;
    case 1: return &launch<kTiles[1].tx, kTiles[1].ty>;
    case 2: return &launch<kTiles[2].tx, kTiles[2].ty>;
    case 3: return &launch<kTiles[3].tx, kTiles[3].ty>;
    default: return nullptr;
  }
}

}  // namespace cato
