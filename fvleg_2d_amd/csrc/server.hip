// Serving form: a request loop with continuous batching.
//
// Requests are queued on submit.  Each tick (1) retires resident solves whose
// device-side time loop finished, copying their results out and returning
// their memory, (2) admits queued requests in arrival order while their HBM
// footprint fits the budget (a first-fit arena allocated once at create, so
// the loop itself never calls hipMalloc), and (3) advances every resident
// solve by `steps_per_tick` steps with one persistent launch per RK stage.
// Solves never interact, so a result is identical whether it ran alone or in
// a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "batch_dispatch.hpp"
#include "runtime.hpp"

namespace cato {
namespace {

struct Arena {
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(size_t n) {
    size = n;
    free_.clear();
    free_[0] = n;
  }
  bool alloc(size_t n, size_t &off) {
    n = (n + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        off = it->first;
        size_t rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~(size_t)255;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  cato_config cfg{};
  int ni = 0, nj = 0;
  std::vector<double> nx, ny, rho, u, v, p;
  double max_time = 0, contour_dt = 0;
  double *out[4] = {};
  double arrival = 0;
  bool done = false;
  double completion = 0;
  long long iters = 0;
  int err = 0;
};

struct Slot {
  bool active = false;
  bool fresh = false;  // admitted this tick: first stage reads the primitives
  int req = -1;
  size_t off = 0, bytes = 0;
  std::unique_ptr<Solve> s;
};

}  // namespace
}  // namespace cato

using namespace cato;

struct cato_server {
  double budget = 0;
  int max_slots = 0, precision = 0, steps_per_tick = 1, tile = 0, nstages = 0;
  hipStream_t stream = nullptr;
  char *arena_mem = nullptr;
  char *table_mem = nullptr;
  size_t table_bytes = 0, loops_off = 0, args_off = 0, variant_off = 0, refs_off = 0, items_off = 0, count_off = 0,
         cursor_off = 0, pass_off = 0;
  int npin = 0;  // resident solves with a +x pressure input (their column passes: [stage slot][npin])
  std::vector<BatchItem> items;  // host copy of the work list (kept until the next tick)
  int item_count = 0;
  int max_items = 0, resident_items = 0;
  Arena arena;
  std::vector<Request> reqs;
  std::deque<int> queue;
  std::vector<Slot> slots;
  BatchLauncher launcher = nullptr;
  int grid = 0;
  int completed = 0;
  int integrator = -1;
  // the resident set of the last table upload; while it is unchanged (no admission, no retirement) the
  // tables stay on the device and the tick's steps replay one captured graph
  bool dirty = true;
  std::vector<int> ids;
  int nitems = 0;
  // resident solves large enough to fill the GPU alone take their own stage-kernel launch per stage (the
  // single-solve kernel, with the registers and LDS of their variant) instead of the merged batch kernel;
  // `tab` is the host copy of the stage arguments those launches pass
  std::vector<int> big;
  std::vector<StageArgs> tab;
  hipGraphExec_t tick_exec = nullptr;
  std::chrono::steady_clock::time_point t0;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int server_tables(cato_server *S) {
  // per-slot device tables: loops, args[(nstages+1) x slots], variant, refs, items
  const int n = S->max_slots;
  size_t off = 0;
  auto a = [](size_t x) { return (x + 255) & ~(size_t)255; };
  S->loops_off = off;
  off += a((size_t)n * sizeof(LoopState));
  S->args_off = off;
  off += a((size_t)(S->nstages + 1) * n * sizeof(StageArgs));
  S->variant_off = off;
  off += a((size_t)n * sizeof(int));
  S->refs_off = off;
  off += a((size_t)n * sizeof(BatchSolveRef));
  S->items_off = off;
  off += a((size_t)S->max_items * sizeof(BatchItem));
  S->count_off = off;
  off += 256;
  S->cursor_off = off;
  off += a((size_t)S->nstages * kNumXcd * sizeof(int));
  S->pass_off = off;
  off += a((size_t)(S->nstages + 1) * n * sizeof(PressurePass));
  S->table_bytes = off;
  int rc = hip_check(hipMalloc(&S->table_mem, S->table_bytes), "server tables");
  if (rc) return rc;
  return hip_check(hipMemset(S->table_mem + S->cursor_off, 0, (size_t)S->nstages * kNumXcd * sizeof(int)),
                   "server cursors");
}

// Rebuild the work list and per-slot stage arguments of the resident set.
// A resident solve of at least this many tiles runs its own stage kernel (one launch fills the GPU).
constexpr int kBigTiles = 2048;

int server_upload(cato_server *S, int &nitems, std::vector<int> &active_ids) {
  const int n = S->max_slots;
  std::vector<StageArgs> &tab = S->tab;
  tab.assign((size_t)(S->nstages + 1) * n, StageArgs{});
  S->big.clear();
  std::vector<int> variant(n, 0);
  std::vector<BatchSolveRef> refs;
  std::vector<BatchItem> &items = S->items;
  items.clear();
  active_ids.clear();
  for (int k = 0; k < n; ++k) {
    Slot &sl = S->slots[k];
    if (!sl.active) continue;
    Solve &sv = *sl.s;
    bool pv = sv.prims_valid;
    sv.prims_valid = false;
    std::vector<StageArgs> st = sv.step_stages(sv.loop, 0.0);
    sv.prims_valid = pv;
    std::vector<StageArgs> first = sv.step_stages(sv.loop, 0.0);
    for (int s = 0; s < S->nstages; ++s) {
      int rc = validate_stage_args(st[s]);
      if (rc) return rc;
      tab[(size_t)s * n + k] = st[s];
    }
    int rc = validate_stage_args(first[0]);
    if (rc) return rc;
    tab[(size_t)S->nstages * n + k] = first[0];
    variant[k] = sv.cfg.flux * LIM_COUNT + sv.cfg.limiter;
    int nt = sv.tiles_x * sv.tiles_y;
    refs.push_back(BatchSolveRef{sv.block_min, sv.block_err, sv.loop, nt, 0});
    active_ids.push_back(k);
  }
  // work list grouped by kernel variant (instruction-cache locality per XCD slice)
  std::vector<int> order(active_ids);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return variant[x] < variant[y]; });
  for (int k : order) {
    const Solve &sv = *S->slots[k].s;
    int nt = sv.tiles_x * sv.tiles_y;
    if (nt >= kBigTiles) {
      S->big.push_back(k);
      continue;
    }
    for (int t = 0; t < nt; ++t) items.push_back({k, t});
  }
  nitems = (int)items.size();
  S->item_count = nitems;
  if (nitems > S->max_items) return fail(CATO_ERR_INVALID, "internal: server work list overflow");
  // +x pressure-input passes of the resident solves that have one, before every stage
  std::vector<int> pin;
  for (int k : active_ids)
    if (S->slots[k].s->cfg.bc_plus_x == BC_PRESSURE_INPUT) pin.push_back(k);
  S->npin = (int)pin.size();
  std::vector<PressurePass> passes((size_t)(S->nstages + 1) * S->npin);
  for (int q = 0; q < S->npin; ++q)
    for (int slot = 0; slot <= S->nstages; ++slot)
      passes[(size_t)slot * S->npin + q] =
          stage_pressure_pass(*S->slots[pin[q]].s, tab[(size_t)slot * n + pin[q]], slot == S->nstages ? 0 : slot);
  hipStream_t st = S->stream;
  int rc;
  if (S->npin && (rc = hip_check(hipMemcpyAsync(S->table_mem + S->pass_off, passes.data(),
                                                passes.size() * sizeof(PressurePass), hipMemcpyHostToDevice, st),
                                 "passes")))
    return rc;
  if ((rc = hip_check(hipMemcpyAsync(S->table_mem + S->args_off, tab.data(), tab.size() * sizeof(StageArgs),
                                     hipMemcpyHostToDevice, st), "args")))
    return rc;
  if ((rc = hip_check(hipMemcpyAsync(S->table_mem + S->variant_off, variant.data(), n * sizeof(int),
                                     hipMemcpyHostToDevice, st), "variant")))
    return rc;
  if (!refs.empty() && (rc = hip_check(hipMemcpyAsync(S->table_mem + S->refs_off, refs.data(),
                                                      refs.size() * sizeof(BatchSolveRef), hipMemcpyHostToDevice, st),
                                       "refs")))
    return rc;
  if (nitems && (rc = hip_check(hipMemcpyAsync(S->table_mem + S->items_off, items.data(),
                                               items.size() * sizeof(BatchItem), hipMemcpyHostToDevice, st),
                                "items")))
    return rc;
  if ((rc = hip_check(hipMemcpyAsync(S->table_mem + S->count_off, &S->item_count, sizeof(int),
                                     hipMemcpyHostToDevice, st), "count")))
    return rc;
  // pageable host vectors: wait before they go out of scope (tab is kept for the big solves' launches)
  return hip_check(hipStreamSynchronize(st), "server upload");
}

int admit_into(cato_server *S, Request &R, Solve &solve, int k, size_t off);

int admit(cato_server *S, int rid) {
  Request &R = S->reqs[rid];
  int k = -1;
  for (int q = 0; q < S->max_slots; ++q)
    if (!S->slots[q].active) {
      k = q;
      break;
    }
  if (k < 0) return 1;  // no slot
  auto sv = std::make_unique<Solve>();
  cato_geom g{R.ni, R.nj, R.nx.data(), R.ny.data()};
  cato_config c = R.cfg;
  c.precision = S->precision;
  c.tile = S->tile;
  int rc = sv->init(c, g);
  if (rc) return rc;
  const int nt = sv->tiles_x * sv->tiles_y;
  size_t bytes = sv->device_bytes(), off = 0;
  if (nt > S->max_items || bytes > S->arena.size) return fail(CATO_ERR_INVALID, "request larger than the server");
  if (S->resident_items + nt > S->max_items) return 1;  // work list full: wait
  if (!S->arena.alloc(bytes, off)) return 1;            // does not fit now: wait
  rc = admit_into(S, R, *sv, k, off);
  if (rc) {
    S->arena.release(off, bytes);
    return rc;
  }
  Slot &sl = S->slots[k];
  sl.active = true;
  sl.fresh = true;
  sl.req = rid;
  sl.off = off;
  sl.bytes = bytes;
  sl.s = std::move(sv);
  S->resident_items += nt;
  return CATO_OK;
}

// Bind a solve to its arena block and upload its state; computes the first dt.
int admit_into(cato_server *S, Request &R, Solve &solve, int k, size_t off) {
  Solve *sv = &solve;
  int rc;
  sv->bind(S->arena_mem + off);
  sv->loop = reinterpret_cast<LoopState *>(S->table_mem + S->loops_off) + k;
  if ((rc = sv->upload_geometry(S->stream))) return rc;
  hipStream_t st = S->stream;
  const std::vector<double> *P[4] = {&R.rho, &R.u, &R.v, &R.p};
  for (int q = 0; q < 4; ++q)
    if ((rc = field_up(st, *sv, sv->P[0][q], P[q]->data(), R.ni, 0, sv->U[1][0]))) return rc;
  FieldGeom fg = sv->fgeom();
  hipLaunchKernelGGL(k_prim_to_cons, dim3((unsigned)((R.ni + 255) / 256), (unsigned)R.nj), dim3(256), 0, st, fg,
                     sv->P[0][0], sv->P[0][1], sv->P[0][2], sv->P[0][3], sv->U[0][0], sv->U[0][1], sv->U[0][2],
                     sv->U[0][3]);
  if ((rc = fill_outflow_column(st, *sv, sv->P[0]))) return rc;
  sv->prims_valid = true;
  LoopState L{};
  double cd = R.contour_dt > 0 ? R.contour_dt : R.max_time;
  L.next_output_time = cd;
  L.contour_dt = cd;
  L.max_time = R.max_time;
  L.max_iterations = 0x7fffffffLL;
  if ((rc = hip_check(hipMemcpyAsync(sv->loop, &L, sizeof(L), hipMemcpyHostToDevice, st), "loop"))) return rc;
  int nb = sv->tiles_x * sv->tiles_y;
  hipLaunchKernelGGL(k_cfl_partial, dim3((unsigned)nb), dim3(kThreads), 0, st, fg, 1, sv->P[0][0], sv->P[0][1],
                     sv->P[0][2], sv->P[0][3], sv->block_min, sv->block_err);
  hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(1024), 0, st, nb, sv->block_min, sv->block_err, sv->result,
                     sv->loop, 1, ConvOut{});
  if ((rc = hip_check(hipGetLastError(), "admit kernels"))) return rc;
  return hip_check(hipStreamSynchronize(st), "admit");  // L lives on this stack frame
}

int retire(cato_server *S, int &retired) {
  retired = 0;
  std::vector<LoopState> H(S->max_slots);
  int rc = hip_check(hipMemcpyAsync(H.data(), S->table_mem + S->loops_off, S->max_slots * sizeof(LoopState),
                                    hipMemcpyDeviceToHost, S->stream),
                     "loops");
  if (rc) return rc;
  if ((rc = hip_check(hipStreamSynchronize(S->stream), "retire"))) return rc;
  for (int k = 0; k < S->max_slots; ++k) {
    Slot &sl = S->slots[k];
    if (!sl.active || !H[k].done) continue;
    Request &R = S->reqs[sl.req];
    Solve &sv = *sl.s;
    for (int q = 0; q < 4; ++q)
      if (R.out[q] && (rc = field_down(S->stream, sv, R.out[q], sv.U[sv.cur][q], R.ni, 0, sv.U[(sv.cur + 1) % 3][0])))
        return rc;
    if ((rc = hip_check(hipStreamSynchronize(S->stream), "result"))) return rc;
    R.done = true;
    R.completion = S->now();
    R.iters = H[k].iteration;
    R.err = H[k].error;
    S->arena.release(sl.off, sl.bytes);
    S->resident_items -= sv.tiles_x * sv.tiles_y;
    sl.active = false;
    sl.s.reset();
    S->completed++;
    retired++;
  }
  return CATO_OK;
}

void drop_tick_graph(cato_server *S) {
  if (S->tick_exec) (void)hipGraphExecDestroy(S->tick_exec);
  S->tick_exec = nullptr;
}

// steps_per_tick steps of the resident set: stage launches (the merged batch kernel over the small solves'
// tiles, one stage-kernel launch per big solve) and the per-solve finalize of every step.  first_fresh:
// stage 0 of the first step takes the table's first-stage slot (fresh solves read their primitives).
int enqueue_tick(cato_server *S, bool first_fresh, int grid, int resident) {
  const int n = S->max_slots;
  const StageArgs *args = reinterpret_cast<const StageArgs *>(S->table_mem + S->args_off);
  for (int step = 0; step < S->steps_per_tick; ++step) {
    for (int s = 0; s < S->nstages; ++s) {
      const int slot = (step == 0 && s == 0 && first_fresh) ? S->nstages : s;
      if (S->npin > 0)
        hipLaunchKernelGGL(k_pressure_columns, dim3((unsigned)S->npin), dim3(1024), 0, S->stream,
                           reinterpret_cast<const PressurePass *>(S->table_mem + S->pass_off) + (size_t)slot * S->npin);
      if (S->nitems > 0) {
        BatchStageArgs b;
        b.args = args + (size_t)slot * n;
        b.variant = reinterpret_cast<const int *>(S->table_mem + S->variant_off);
        b.items = reinterpret_cast<const BatchItem *>(S->table_mem + S->items_off);
        b.count = reinterpret_cast<const int *>(S->table_mem + S->count_off);
        b.cursor = reinterpret_cast<int *>(S->table_mem + S->cursor_off) + (size_t)s * kNumXcd;
        S->launcher(dim3((unsigned)grid), S->stream, b);
      }
      for (int k : S->big) {
        const Solve &sv = *S->slots[k].s;
        const StageArgs &a = S->tab[(size_t)slot * n + k];
        sv.launcher(dim3((unsigned)(sv.tiles_x * sv.tiles_y)), S->stream, a);
        // the merged kernel writes a +x outflow ghost column in its epilogue; the stage kernel leaves it to this
        if (S->precision == CATO_PRECISION_FP64) {
          int rc = fill_outflow_column(S->stream, sv, a.prim_out);
          if (rc) return rc;
        }
      }
    }
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)resident), dim3(kThreads), 0, S->stream,
                       reinterpret_cast<const BatchSolveRef *>(S->table_mem + S->refs_off), 2,
                       reinterpret_cast<int *>(S->table_mem + S->cursor_off), S->nstages * kNumXcd);
  }
  return hip_check(hipGetLastError(), "server step");
}

}  // namespace

extern "C" {

int cato_server_create(double hbm_budget_bytes, int max_slots, int precision, int steps_per_tick,
                       cato_server **out) {
  if (!out) return fail(CATO_ERR_INVALID, "null output");
  *out = nullptr;
  if (!(hbm_budget_bytes > 0) || max_slots <= 0 || steps_per_tick <= 0)
    return fail(CATO_ERR_INVALID, "bad server parameters");
  if (precision != CATO_PRECISION_FP64 && precision != CATO_PRECISION_FP32_FLUX)
    return fail(CATO_ERR_INVALID, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CATO_ERR_NO_DEVICE, "no HIP device");
  cato_server *S = new cato_server();
  S->budget = hbm_budget_bytes;
  S->max_slots = max_slots;
  S->precision = precision;
  S->steps_per_tick = steps_per_tick;
  S->tile = 0;
  S->t0 = std::chrono::steady_clock::now();
  S->slots.resize(max_slots);
  S->launcher = batch_launcher(precision, S->tile);
  // the largest tile count that fits the budget bounds the work list
  // work-list capacity: a solve holds at least 128 B of device state per cell and a
  // 32 x 8 tile covers up to 256 cells; partial tiles at the edges add the slack
  S->max_items = (int)std::min<double>(2.0 * hbm_budget_bytes / (128.0 * 256.0) + 4096.0 * max_slots, 1 << 26);
  int rc = hip_check(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking), "stream");
  if (!rc) rc = hip_check(hipMalloc(&S->arena_mem, (size_t)hbm_budget_bytes), "server arena");
  if (rc) {
    cato_server_destroy(S);
    return rc;
  }
  S->arena.init((size_t)hbm_budget_bytes);
  hipDeviceProp_t prop;
  int dev = 0;
  (void)hipGetDevice(&dev);
  int cus = (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
                ? prop.multiProcessorCount
                : 256;
  S->grid = 4 * cus;
  *out = S;
  return CATO_OK;
}

int cato_server_destroy(cato_server *S) {
  if (!S) return CATO_OK;
  if (S->stream) (void)hipStreamSynchronize(S->stream);
  drop_tick_graph(S);
  S->slots.clear();
  if (S->arena_mem) (void)hipFree(S->arena_mem);
  if (S->table_mem) (void)hipFree(S->table_mem);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
  return CATO_OK;
}

int cato_server_submit(cato_server *S, const cato_request *req, double arrival_time) {
  if (!S || !req) return fail(CATO_ERR_INVALID, "null argument");
  const cato_geom &g = req->geom;
  int nh = req->cfg.n_ghost_layers;
  if (g.ni <= 0 || g.nj <= 0 || nh <= 0 || !g.node_x || !g.node_y || !req->rho || !req->u || !req->v || !req->p)
    return fail(CATO_ERR_INVALID, "incomplete request");
  if (req->cfg.bc_plus_x == BC_PRESSURE_INPUT) {  // the constant drive (a request carries no table)
    Solve probe;
    probe.cfg = req->cfg;
    int rc = check_pressure_constants(probe);
    if (rc) return rc;
  }
  if (S->integrator < 0) {
    S->integrator = req->cfg.integrator;
    Solve probe;
    probe.cfg = req->cfg;
    probe.prims_valid = false;
    std::vector<StageArgs> st = probe.step_stages(nullptr, 0.0);
    S->nstages = (int)st.size();
    int rc = server_tables(S);
    if (rc) return rc;
  } else if (req->cfg.integrator != S->integrator) {
    return fail(CATO_ERR_UNSUPPORTED, "all requests of a server share one integrator");
  }
  Request R;
  R.cfg = req->cfg;
  R.ni = g.ni;
  R.nj = g.nj;
  size_t nn = (size_t)(g.ni + 2 * nh + 1) * (g.nj + 2 * nh + 1), nc = (size_t)g.ni * g.nj;
  R.nx.assign(g.node_x, g.node_x + nn);
  R.ny.assign(g.node_y, g.node_y + nn);
  R.rho.assign(req->rho, req->rho + nc);
  R.u.assign(req->u, req->u + nc);
  R.v.assign(req->v, req->v + nc);
  R.p.assign(req->p, req->p + nc);
  R.max_time = req->max_time;
  R.contour_dt = req->contour_interval_dt;
  R.out[0] = req->out_rho;
  R.out[1] = req->out_rho_u;
  R.out[2] = req->out_rho_v;
  R.out[3] = req->out_rho_E;
  R.arrival = arrival_time;
  S->reqs.push_back(std::move(R));
  int id = (int)S->reqs.size() - 1;
  S->queue.push_back(id);
  return id;
}

int cato_server_tick(cato_server *S, int *n_resident, int *n_queued, int *n_completed_total) {
  if (!S) return fail(CATO_ERR_INVALID, "null server");
  int rc, retired = 0;
  if (S->table_mem && (rc = retire(S, retired))) return rc;
  if (retired) S->dirty = true;
  // admission in arrival order while slots and memory allow
  while (!S->queue.empty()) {
    int rid = S->queue.front();
    rc = admit(S, rid);
    if (rc == 1) break;
    if (rc) {
      S->reqs[rid].done = true;
      S->reqs[rid].err = -1;
      S->reqs[rid].completion = S->now();
      S->queue.pop_front();
      S->completed++;
      continue;
    }
    S->queue.pop_front();
    S->dirty = true;
  }
  if (S->table_mem && S->dirty) {
    drop_tick_graph(S);
    if ((rc = server_upload(S, S->nitems, S->ids))) return rc;
    // work cursors start at zero (the finalize kernel resets them after every step)
    if ((rc = hip_check(hipMemsetAsync(S->table_mem + S->cursor_off, 0, (size_t)S->nstages * kNumXcd * sizeof(int),
                                       S->stream), "cursors")))
      return rc;
    S->dirty = false;
  }
  const int resident = S->table_mem ? (int)S->ids.size() : 0;
  if (resident > 0) {
    // slot `nstages` of the table holds stage 0 as the first step of this tick
    // runs it: fresh solves read their uploaded primitives, the others their state
    bool any_fresh = false;
    for (int k : S->ids) any_fresh = any_fresh || S->slots[k].fresh;
    const int grid = std::max(kNumXcd, (std::min(S->nitems, S->grid) / kNumXcd) * kNumXcd);
    if (any_fresh) {
      if ((rc = enqueue_tick(S, true, grid, resident))) return rc;
    } else {
      // the same launches every tick until the resident set changes: one graph
      if (!S->tick_exec) {
        if ((rc = hip_check(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal), "capture tick")))
          return rc;
        rc = enqueue_tick(S, false, grid, resident);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(S->stream, &g);
        if (!rc && e == hipSuccess) e = hipGraphInstantiate(&S->tick_exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        if ((rc = hip_check(e, "tick graph"))) {
          S->tick_exec = nullptr;
          return rc;
        }
      }
      if ((rc = hip_check(hipGraphLaunch(S->tick_exec, S->stream), "tick graph launch"))) return rc;
    }
    for (int k : S->ids) {
      S->slots[k].fresh = false;
      S->slots[k].s->prims_valid = false;
    }
  }
  if (n_resident) *n_resident = resident;
  if (n_queued) *n_queued = (int)S->queue.size();
  if (n_completed_total) *n_completed_total = S->completed;
  return CATO_OK;
}

int cato_server_poll(cato_server *S, int id, int *done, double *completion_time, long long *iterations,
                     int *error_code) {
  if (!S) return fail(CATO_ERR_INVALID, "null server");
  if (id < 0 || id >= (int)S->reqs.size()) return fail(CATO_ERR_INVALID, "unknown request id");
  const Request &R = S->reqs[id];
  if (done) *done = R.done ? 1 : 0;
  if (completion_time) *completion_time = R.completion;
  if (iterations) *iterations = R.iters;
  if (error_code) *error_code = R.err;
  return CATO_OK;
}

double cato_server_clock(cato_server *S) { return S ? S->now() : 0.0; }

}  // extern "C"
