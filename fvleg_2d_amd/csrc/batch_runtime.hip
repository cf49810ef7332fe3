// Batched form of the CATO hot path: many independent solves resident in one
// device allocation and advanced together, one persistent launch per RK stage
// plus one per-solve finalize.  Steady-state steps are captured once into a
// hipGraph and replayed; dt, time and the done flag of every solve live on the
// device, so the host only polls every few steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "batch_dispatch.hpp"
#include "batch_runtime.hpp"
#include "runtime.hpp"

namespace cato {

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

#define BCK(expr)                                 \
  do {                                            \
    int _rc = hip_check((expr), #expr);           \
    if (_rc != CATO_OK) return _rc;               \
  } while (0)

int BatchEngine::init(int n, const cato_config *cfgs, const cato_geom *geoms, int prec, int tile_id,
                      hipStream_t s) {
  if (n <= 0 || !cfgs || !geoms) return fail(CATO_ERR_INVALID, "empty batch");
  stream = s;
  precision = prec;
  tile = tile_id < 0 ? 0 : tile_id;
  if (tile >= kNumTiles) return fail(CATO_ERR_INVALID, "unknown tile id");
  launcher = batch_launcher(precision, tile);
  if (!launcher) return fail(CATO_ERR_INVALID, "no batch kernel for this precision/tile");
  solves.resize(n);
  int integ = cfgs[0].integrator;
  for (int k = 0; k < n; ++k) {
    cato_config c = cfgs[k];
    c.precision = precision;
    c.tile = tile;
    if (c.integrator != integ) return fail(CATO_ERR_UNSUPPORTED, "all solves of a batch share one integrator");
    int rc = solves[k].init(c, geoms[k]);
    if (rc) return rc;
    if (c.bc_plus_x == BC_PRESSURE_INPUT) pin.push_back(k);
  }
  // one allocation: per-solve regions, then shared tables
  offsets.resize(n);
  size_t off = 0;
  for (int k = 0; k < n; ++k) {
    offsets[k] = off;
    off += align256(solves[k].device_bytes());
  }
  nitems = 0;
  for (auto &sv : solves) nitems += sv.tiles_x * sv.tiles_y;
  nstages = (int)solves[0].step_stages(nullptr, 0.0).size();
  loops_off = off;
  off += align256((size_t)n * sizeof(LoopState));
  args_off = off;
  off += align256((size_t)(nstages + 1) * n * sizeof(StageArgs));
  variant_off = off;
  off += align256((size_t)n * sizeof(int));
  items_off = off;
  off += align256((size_t)nitems * sizeof(BatchItem));
  refs_off = off;
  off += align256((size_t)n * sizeof(BatchSolveRef));
  count_off = off;
  off += 256;
  cursor_off = off;
  off += align256((size_t)nstages * kNumXcd * sizeof(int));
  pass_off = off;
  off += align256((size_t)(nstages + 1) * pin.size() * sizeof(PressurePass));
  total_bytes = off;
  BCK(hipMalloc(&mem, total_bytes));
  for (int k = 0; k < n; ++k) {
    solves[k].bind(mem + offsets[k]);
    solves[k].loop = reinterpret_cast<LoopState *>(mem + loops_off) + k;
    int rc = solves[k].upload_geometry(stream);
    if (rc) return rc;
  }
  // work list grouped by solve, variants, finalize references
  std::vector<BatchItem> &items = live_items;
  items.clear();
  items.reserve(nitems);
  live_solves = n;
  std::vector<int> variant(n);
  std::vector<BatchSolveRef> refs(n);
  for (int k = 0; k < n; ++k) {
    const Solve &sv = solves[k];
    int nt = sv.tiles_x * sv.tiles_y;
    variant[k] = sv.cfg.flux * LIM_COUNT + sv.cfg.limiter;
    refs[k] = BatchSolveRef{sv.block_min, sv.block_err, sv.loop, nt, 0};
  }
  // work list grouped by variant: each XCD slice then runs few kernel variants,
  // which keeps the instruction cache warm
  order.resize(n);
  for (int k = 0; k < n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return variant[x] < variant[y]; });
  for (int k : order) {
    int nt = solves[k].tiles_x * solves[k].tiles_y;
    for (int t = 0; t < nt; ++t) items.push_back({k, t});
  }
  BCK(hipMemcpyAsync(mem + variant_off, variant.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
  BCK(hipMemcpyAsync(mem + items_off, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice,
                     stream));
  BCK(hipMemcpyAsync(mem + refs_off, refs.data(), n * sizeof(BatchSolveRef), hipMemcpyHostToDevice, stream));
  BCK(hipMemcpyAsync(mem + count_off, &nitems, sizeof(int), hipMemcpyHostToDevice, stream));
  BCK(hipMemsetAsync(mem + cursor_off, 0, (size_t)nstages * kNumXcd * sizeof(int), stream));
  hipDeviceProp_t prop;
  int dev = 0;
  BCK(hipGetDevice(&dev));
  BCK(hipGetDeviceProperties(&prop, dev));
  int cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  grid = std::min(nitems, 4 * cus);
  grid = std::max(kNumXcd, (grid / kNumXcd) * kNumXcd);
  return hip_check(hipStreamSynchronize(stream), "batch init");
}

void BatchEngine::release() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  if (mem) (void)hipFree(mem);
  mem = nullptr;
  for (auto &sv : solves) {
    if (sv.pi_tab) (void)hipFree(sv.pi_tab);
    sv.pi_tab = nullptr;
    sv.pi_n = 0;
  }
}

// pressure_input_bc_t's table of solve k (times, pressures, densities; n = 0: the constant drive).
int BatchEngine::set_pressure_input(int k, const double *time, const double *pressure, const double *density,
                                    int n) {
  Solve &sv = solves[k];
  if (sv.cfg.bc_plus_x != BC_PRESSURE_INPUT) return fail(CATO_ERR_INVALID, "no pressure_input boundary at +x");
  int rc = check_pressure_table(time, pressure, density, n);
  if (rc) return rc;
  BCK(hipStreamSynchronize(stream));
  if (sv.pi_tab) (void)hipFree(sv.pi_tab);
  sv.pi_tab = nullptr;
  sv.pi_n = 0;
  if (n == 0) return CATO_OK;
  std::vector<double> tab((size_t)3 * n);
  std::copy(time, time + n, tab.begin());
  std::copy(pressure, pressure + n, tab.begin() + n);
  std::copy(density, density + n, tab.begin() + 2 * n);
  BCK(hipMalloc((void **)&sv.pi_tab, tab.size() * sizeof(double)));
  sv.pi_n = n;
  BCK(hipMemcpyAsync(sv.pi_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  return hip_check(hipStreamSynchronize(stream), "batch set_pressure_input");
}

// Stage arguments of every solve: table[stage][solve]; slot `nstages` holds the
// first stage of the first step (primitives from the caller's arrays).
int BatchEngine::upload_args() {
  const int n = (int)solves.size();
  std::vector<StageArgs> tab((size_t)(nstages + 1) * n);
  for (int k = 0; k < n; ++k) {
    Solve &sv = solves[k];
    bool pv = sv.prims_valid;
    sv.prims_valid = false;
    std::vector<StageArgs> st = sv.step_stages(sv.loop, 0.0);
    sv.prims_valid = pv;
    for (int s = 0; s < nstages; ++s) {
      int rc = validate_stage_args(st[s]);
      if (rc) return rc;
      tab[(size_t)s * n + k] = st[s];
    }
    std::vector<StageArgs> first = sv.step_stages(sv.loop, 0.0);
    int rc = validate_stage_args(first[0]);
    if (rc) return rc;
    tab[(size_t)nstages * n + k] = first[0];
  }
  BCK(hipMemcpyAsync(mem + args_off, tab.data(), tab.size() * sizeof(StageArgs), hipMemcpyHostToDevice, stream));
  // the pressure-input passes before each stage, from the same stage arguments
  const size_t np = pin.size();
  std::vector<PressurePass> passes((size_t)(nstages + 1) * np);
  for (size_t q = 0; q < np; ++q) {
    const int k = pin[q];
    int rc = check_pressure_constants(solves[k]);
    if (rc) return rc;
    for (int slot = 0; slot <= nstages; ++slot)
      passes[(size_t)slot * np + q] = stage_pressure_pass(solves[k], tab[(size_t)slot * n + k], slot == nstages ? 0 : slot);
  }
  if (np)
    BCK(hipMemcpyAsync(mem + pass_off, passes.data(), passes.size() * sizeof(PressurePass), hipMemcpyHostToDevice,
                       stream));
  return hip_check(hipStreamSynchronize(stream), "batch args");  // tab and passes are pageable and local
}

int BatchEngine::launch_step(bool first) {
  const int n = (int)solves.size();
  const StageArgs *args = reinterpret_cast<const StageArgs *>(mem + args_off);
  for (int s = 0; s < nstages; ++s) {
    const int slot = (first && s == 0) ? nstages : s;
    if (!pin.empty()) {
      hipLaunchKernelGGL(k_pressure_columns, dim3((unsigned)pin.size()), dim3(1024), 0, stream,
                         reinterpret_cast<const PressurePass *>(mem + pass_off) + (size_t)slot * pin.size());
      BCK(hipGetLastError());
    }
    BatchStageArgs b;
    b.args = args + (size_t)slot * n;
    b.variant = reinterpret_cast<const int *>(mem + variant_off);
    b.items = reinterpret_cast<const BatchItem *>(mem + items_off);
    b.count = reinterpret_cast<const int *>(mem + count_off);
    b.cursor = reinterpret_cast<int *>(mem + cursor_off) + (size_t)s * kNumXcd;
    launcher(dim3((unsigned)grid), stream, b);
    BCK(hipGetLastError());
  }
  hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)n), dim3(kThreads), 0, stream,
                     reinterpret_cast<const BatchSolveRef *>(mem + refs_off), 2, reinterpret_cast<int *>(mem + cursor_off),
                     nstages * kNumXcd);
  BCK(hipGetLastError());
  return CATO_OK;
}

// Drop finished solves from the work list so the persistent grid (split in
// contiguous eighths over the XCDs) stays balanced as solves end at different
// steps.  Called after a stream synchronisation; the copies are ordered before
// the next replay on the same stream.
int BatchEngine::compact(const std::vector<LoopState> &H) {
  const int n = (int)solves.size();
  int live = 0;
  for (int k = 0; k < n; ++k) live += H[k].done ? 0 : 1;
  if (live == live_solves) return CATO_OK;
  live_solves = live;
  live_items.clear();
  for (int k : order) {
    if (H[k].done) continue;
    int nt = solves[k].tiles_x * solves[k].tiles_y;
    for (int t = 0; t < nt; ++t) live_items.push_back({k, t});
  }
  live_count = (int)live_items.size();
  if (live_count)
    BCK(hipMemcpyAsync(mem + items_off, live_items.data(), live_items.size() * sizeof(BatchItem),
                       hipMemcpyHostToDevice, stream));
  BCK(hipMemcpyAsync(mem + count_off, &live_count, sizeof(int), hipMemcpyHostToDevice, stream));
  return hip_check(hipStreamSynchronize(stream), "batch compact");
}

int BatchEngine::run(const double *max_time, const double *contour_dt, long long max_iterations, int use_graph,
                     long long *iters, double *times, int *errors) {
  const int n = (int)solves.size();
  if (max_iterations <= 0) max_iterations = 0x7fffffffLL;
  std::vector<LoopState> L(n);
  for (int k = 0; k < n; ++k) {
    L[k] = LoopState{};
    double cd = (contour_dt && contour_dt[k] > 0.0) ? contour_dt[k] : max_time[k];
    L[k].next_output_time = 0.0 + cd;
    L[k].contour_dt = cd;
    L[k].max_time = max_time[k];
    L[k].max_iterations = max_iterations;
  }
  BCK(hipMemcpyAsync(mem + loops_off, L.data(), n * sizeof(LoopState), hipMemcpyHostToDevice, stream));
  BCK(hipMemsetAsync(mem + cursor_off, 0, (size_t)nstages * kNumXcd * sizeof(int), stream));  // outside any graph
  int rc = upload_args();  // synchronises, so L may go
  if (rc) return rc;
  if ((rc = compact(L))) return rc;  // every solve live again (a batch may be run more than once)
  // first dt of every solve
  for (int k = 0; k < n; ++k) {
    Solve &sv = solves[k];
    if ((rc = refresh_prims(stream, sv))) return rc;  // a precision-form batch that already ran
    FieldGeom g = sv.fgeom();
    int nb = sv.tiles_x * sv.tiles_y;
    const double *a[4];
    for (int v = 0; v < 4; ++v) a[v] = sv.P[0][v];  // primitives of U^n
    hipLaunchKernelGGL(k_cfl_partial, dim3((unsigned)nb), dim3(kThreads), 0, stream, g, 1, a[0],
                       a[1], a[2], a[3], sv.block_min, sv.block_err);
    hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(1024), 0, stream, nb, sv.block_min, sv.block_err, sv.result,
                       sv.loop, 1, ConvOut{});
    BCK(hipGetLastError());
  }
  // first step eagerly (its first stage reads the caller's primitives)
  if ((rc = launch_step(true))) return rc;
  for (auto &sv : solves) {
    sv.prims_valid = false;
    sv.prims_stale = sv.cfg.precision != CATO_PRECISION_FP64;  // no primitives stored by those stages
  }
  if (use_graph && !exec) {
    BCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    rc = launch_step(false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    BCK(e);
    graph = g;
    BCK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  std::vector<LoopState> H(n);
  const int chunk = 16;
  while (true) {
    for (int c = 0; c < chunk; ++c) {
      if (use_graph) {
        BCK(hipGraphLaunch(exec, stream));
      } else if ((rc = launch_step(false))) {
        return rc;
      }
    }
    BCK(hipMemcpyAsync(H.data(), mem + loops_off, n * sizeof(LoopState), hipMemcpyDeviceToHost, stream));
    BCK(hipStreamSynchronize(stream));
    bool all = true;
    for (int k = 0; k < n; ++k) all = all && H[k].done;
    if (all) break;
    if ((rc = compact(H))) return rc;
  }
  for (int k = 0; k < n; ++k) {
    if (iters) iters[k] = H[k].iteration;
    if (times) times[k] = H[k].time;
    if (errors) errors[k] = H[k].error;
  }
  return CATO_OK;
}

}  // namespace cato

using namespace cato;

struct cato_batch {
  BatchEngine e;
  hipStream_t stream = nullptr;
};

extern "C" {

int cato_batch_create(int n, const cato_config *cfgs, const cato_geom *geoms, int precision, int tile,
                      cato_batch **out) {
  if (!out) return fail(CATO_ERR_INVALID, "null output");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CATO_ERR_NO_DEVICE, "no HIP device");
  if (precision != CATO_PRECISION_FP64 && precision != CATO_PRECISION_FP32_FLUX)
    return fail(CATO_ERR_INVALID, "unknown precision");
  cato_batch *b = new cato_batch();
  int rc = hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
  if (!rc) rc = b->e.init(n, cfgs, geoms, precision, tile, b->stream);
  if (rc) {
    cato_batch_destroy(b);
    return rc;
  }
  *out = b;
  return CATO_OK;
}

int cato_batch_destroy(cato_batch *b) {
  if (!b) return CATO_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->e.release();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CATO_OK;
}

static int batch_solve(cato_batch *b, int k, Solve **s) {
  if (!b) return fail(CATO_ERR_INVALID, "null batch");
  if (k < 0 || k >= (int)b->e.solves.size()) return fail(CATO_ERR_INVALID, "solve index out of range");
  *s = &b->e.solves[k];
  return CATO_OK;
}

int cato_batch_set_primitive(cato_batch *b, int k, const double *rho, const double *u, const double *v,
                             const double *p, int ld, int halo) {
  Solve *s;
  int rc = batch_solve(b, k, &s);
  if (rc) return rc;
  const double *P[4] = {rho, u, v, p};
  double *stg = s->U[(s->cur + 1) % 3][0];
  for (int q = 0; q < 4; ++q)
    if ((rc = field_up(b->stream, *s, s->P[0][q], P[q], ld, halo, stg))) return rc;
  FieldGeom g = s->fgeom();
  hipLaunchKernelGGL(k_prim_to_cons, dim3((unsigned)((s->ni + 255) / 256), (unsigned)s->nj), dim3(256), 0, b->stream,
                     g, s->P[0][0], s->P[0][1], s->P[0][2], s->P[0][3], s->U[s->cur][0], s->U[s->cur][1],
                     s->U[s->cur][2], s->U[s->cur][3]);
  if ((rc = hip_check(hipGetLastError(), "prim_to_cons"))) return rc;
  if ((rc = fill_outflow_column(b->stream, *s, s->P[0]))) return rc;
  s->prims_valid = true;
  s->prims_stale = false;
  return hip_check(hipStreamSynchronize(b->stream), "batch set_primitive");
}

int cato_batch_set_pressure_input(cato_batch *b, int k, const double *time, const double *pressure,
                                  const double *density, int n) {
  Solve *s;
  int rc = batch_solve(b, k, &s);
  if (rc) return rc;
  return b->e.set_pressure_input(k, time, pressure, density, n);
}

int cato_batch_get_conserved(cato_batch *b, int k, double *rho, double *rho_u, double *rho_v, double *rho_E, int ld,
                             int halo) {
  Solve *s;
  int rc = batch_solve(b, k, &s);
  if (rc) return rc;
  double *H[4] = {rho, rho_u, rho_v, rho_E};
  double *stg = s->U[(s->cur + 1) % 3][0];
  for (int q = 0; q < 4; ++q)
    if (H[q] && (rc = field_down(b->stream, *s, H[q], s->U[s->cur][q], ld, halo, stg))) return rc;
  return hip_check(hipStreamSynchronize(b->stream), "batch get_conserved");
}

int cato_batch_run(cato_batch *b, const double *max_time, const double *contour_interval_dt, long long max_iterations,
                   int use_graph, long long *iterations, double *times, int *errors) {
  if (!b || !max_time) return fail(CATO_ERR_INVALID, "null argument");
  return b->e.run(max_time, contour_interval_dt, max_iterations, use_graph, iterations, times, errors);
}

}  // extern "C"
