// Batched form: many independent solves (any grid size, flux, limiter) advanced
// by ONE persistent launch per RK stage.  Each workgroup walks a contiguous
// slice of the (solve, tile) work list of its XCD and dispatches the solve's
// (flux, limiter) variant of the fused stage tile; all variants share one LDS
// buffer sized for the largest.  Solves whose time loop is done are skipped,
// and the host drops them from the work list at its next poll.
#pragma once

#include "aux_kernels.hpp"

namespace cato {

struct BatchItem {
  int solve;
  int tile;
};

struct BatchSolveRef {
  double *block_min;
  int *block_err;
  LoopState *loop;
  int ntiles;
  int pad;
};

struct BatchStageArgs {
  const StageArgs *args;   // per solve, this stage
  const int *variant;      // per solve: flux * LIM_COUNT + limiter
  const BatchItem *items;  // work list, grouped by variant and solve
  const int *count;        // number of live entries of items (device-resident: graph replays see updates)
  int *cursor;             // kNumXcd work cursors of this stage: zero at launch, reset by the finalize kernel
};

template <typename R, int TX, int TY> constexpr size_t batch_smem_bytes() {
  size_t a = sizeof(TileSmem<LIM_MINMOD, R, TX, TY>);
  size_t b = sizeof(TileSmem<LIM_MLP5, R, TX, TY>);
  return a > b ? a : b;
}

#define CATO_BATCH_CASE(F, L)                                                                                 \
  case (F) * LIM_COUNT + (L):                                                                                 \
    stage_tile<F, L, R, TX, TY, true>(a, it.tile, dt, *reinterpret_cast<TileSmem<L, R, TX, TY> *>(smem), red); \
    break;
#define CATO_BATCH_FLUX(F)   \
  CATO_BATCH_CASE(F, 0)      \
  CATO_BATCH_CASE(F, 1)      \
  CATO_BATCH_CASE(F, 2)      \
  CATO_BATCH_CASE(F, 3)      \
  CATO_BATCH_CASE(F, 4)

// Waves per SIMD the LDS footprint allows (at most 4): the register budget is
// capped to match, since the forty variants merged in one body would
// otherwise claim far more registers than any single variant needs.
template <typename R, int TX, int TY> constexpr int batch_waves_per_simd() {
  constexpr size_t lds = batch_smem_bytes<R, TX, TY>() + sizeof(RedSmem);
  constexpr int fit = (int)((160 * 1024) / lds);
  return fit < 1 ? 1 : (fit > 4 ? 4 : fit);
}

template <typename R, int TX, int TY>
__global__ __launch_bounds__(kThreads, (batch_waves_per_simd<R, TX, TY>())) void batch_stage_kernel(BatchStageArgs b) {
  __shared__ __attribute__((aligned(16))) char smem[batch_smem_bytes<R, TX, TY>()];
  __shared__ RedSmem red;
  __shared__ int s_item;
  // The work list is cut into kNumXcd contiguous slices.  Workgroups of XCD x
  // take items of slice x through its cursor (neighbouring tiles stay in one
  // L2), then help the other slices: items differ in cost by variant, so a
  // static split would leave XCDs idle while others finish.
  const int nitems = *b.count;
  const int x0 = blockIdx.x % kNumXcd;
  for (int k = 0; k < kNumXcd; ++k) {
    const int x = (x0 + k) % kNumXcd;
    const int lo = (int)(((long long)nitems * x) / kNumXcd), hi = (int)(((long long)nitems * (x + 1)) / kNumXcd);
    while (true) {
      __syncthreads();  // the previous item (its LDS and s_item) is finished
      if (threadIdx.x == 0) {
        const int w = lo + atomicAdd(&b.cursor[x], 1);
        s_item = (w < hi) ? w : -1;
      }
      __syncthreads();
      const int w = s_item;
      if (w < 0) break;
      const BatchItem it = b.items[w];
      const StageArgs &a = b.args[it.solve];
      if (a.loop && a.loop->done) continue;
      const double dt = a.loop ? a.loop->dt : a.dt_value;
      switch (b.variant[it.solve]) {
        CATO_BATCH_FLUX(0)
        CATO_BATCH_FLUX(1)
        CATO_BATCH_FLUX(2)
        CATO_BATCH_FLUX(3)
        CATO_BATCH_FLUX(4)
        CATO_BATCH_FLUX(5)
        CATO_BATCH_FLUX(6)
        CATO_BATCH_FLUX(7)
        CATO_BATCH_FLUX(8)
        default: break;
      }
    }
  }
}

#undef CATO_BATCH_CASE
#undef CATO_BATCH_FLUX

// Per-solve end-of-step reduction: one workgroup per solve.  Workgroup 0 also
// zeroes the stage work cursors for the next step (all stages have finished).
static __global__ __launch_bounds__(kThreads) void batch_finalize_kernel(const BatchSolveRef *refs, int mode,
                                                                       int *cursors, int ncursors) {
  __shared__ double sm[kThreads / 64];
  __shared__ int se[kThreads / 64];
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < ncursors; k += kThreads) cursors[k] = 0;
  const BatchSolveRef r = refs[blockIdx.x];
  LoopState *L = r.loop;
  if (mode == 2 && L->done) return;
  double m = __builtin_inf();
  int e = 0;
  for (int k = threadIdx.x; k < r.ntiles; k += kThreads) {
    double v = r.block_min[k];
    m = (v < m) ? v : m;
    e |= r.block_err[k];
    r.block_err[k] = 0;
  }
  m = wave_min(m);
  e = wave_or(e);
  if ((threadIdx.x & 63) == 0) {
    sm[threadIdx.x >> 6] = m;
    se[threadIdx.x >> 6] = e;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 0; k < kThreads / 64; ++k) {
    m = (sm[k] < m) ? sm[k] : m;
    e |= se[k];
  }
  loop_update(L, m, e, mode);
}

}  // namespace cato
