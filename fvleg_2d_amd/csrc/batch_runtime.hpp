// Batched-form engine shared by the cato_batch C API and the serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "batch_dispatch.hpp"
#include "runtime.hpp"

namespace cato {

struct BatchEngine {
  std::vector<Solve> solves;
  std::vector<size_t> offsets;
  hipStream_t stream = nullptr;
  char *mem = nullptr;
  size_t total_bytes = 0;
  size_t loops_off = 0, args_off = 0, variant_off = 0, items_off = 0, refs_off = 0, count_off = 0, cursor_off = 0;
  // +x pressure input: the solves that have one, and their column passes [stage slot][solve of `pin`]
  std::vector<int> pin;
  size_t pass_off = 0;
  std::vector<BatchItem> live_items;  // host copy of the current work list
  std::vector<int> order;             // solves ordered by kernel variant
  int live_solves = 0, live_count = 0;
  int precision = 0, tile = 0, nstages = 0, nitems = 0, grid = 0;
  BatchLauncher launcher = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  int init(int n, const cato_config *cfgs, const cato_geom *geoms, int prec, int tile_id, hipStream_t s);
  void release();
  int upload_args();
  int launch_step(bool first);
  int compact(const std::vector<LoopState> &H);
  int set_pressure_input(int k, const double *time, const double *pressure, const double *density, int n);
  int run(const double *max_time, const double *contour_dt, long long max_iterations, int use_graph,
          long long *iters, double *times, int *errors);
};

}  // namespace cato
