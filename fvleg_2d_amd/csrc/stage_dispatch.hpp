// Launch table of the fused stage kernel: one entry per (flux, limiter,
// precision, tile shape).  Each flux family is instantiated in its own
// translation unit (stage_f<id>.hip) so the library builds in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "cato_types.hpp"

namespace cato {

struct TileShape {
  int tx, ty;
};
constexpr int kNumTiles = 4;
constexpr TileShape kTiles[kNumTiles] = {{32, 8}, {32, 16}, {16, 16}, {64, 8}};
// the one tile shape instantiated for general quadrilateral cells
constexpr int kQuadTile = 2;

using StageLauncher = void (*)(dim3 grid, hipStream_t stream, const StageArgs &args);

StageLauncher stage_launcher_f0(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f1(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f2(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f3(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f4(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f5(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f6(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f7(int lim, int prec, int tile, int quad);
StageLauncher stage_launcher_f8(int lim, int prec, int tile, int quad);

inline StageLauncher stage_launcher(int flux, int lim, int prec, int tile, int quad) {
  switch (flux) {
    case 0: return stage_launcher_f0(lim, prec, tile, quad);
    case 1: return stage_launcher_f1(lim, prec, tile, quad);
    case 2: return stage_launcher_f2(lim, prec, tile, quad);
    case 3: return stage_launcher_f3(lim, prec, tile, quad);
    case 4: return stage_launcher_f4(lim, prec, tile, quad);
    case 5: return stage_launcher_f5(lim, prec, tile, quad);
    case 6: return stage_launcher_f6(lim, prec, tile, quad);
    case 7: return stage_launcher_f7(lim, prec, tile, quad);
    case 8: return stage_launcher_f8(lim, prec, tile, quad);
    default: return nullptr;
  }
}

}  // namespace cato
