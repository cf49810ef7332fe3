// One solve split into px x py sub-domains on one GPU: the shape a run of one
// large solve over several GPUs takes (the reference decomposes the grid into
// near-square blocks per coarray image, mod_parallel.f90:41-124, and exchanges
// nh-deep edge and corner halos of every field, field_2d.f90:551-654 with the
// shapes of 589-597).  Every sub-domain is an ordinary solve whose inner sides
// (and periodic sides wrapped onto another sub-domain) take their ghost cells
// from a halo store.  After every RK stage the primitives the next stage reads
// are exchanged:
//   pack   (k_halo_pack): each sub-domain gathers the nh columns / rows / corner
//          blocks each neighbour needs from its primitive arrays into a send
//          buffer, already in the receiver's halo-store layout;
//   unpack (k_halo_unpack): each sub-domain pulls its neighbours' send regions
//          into its halo store (over xGMI this is a one-sided read of a peer's
//          send buffer, or the landing zone of a send/recv pair).
// The CFL dt of a step is the minimum over the sub-domains (the 8-byte
// all-reduce of mod_fluid.f90:560).  The exchange copies values, and the halo
// store is read exactly where the unsplit solve reads its neighbours' cells, so
// the split solve gives the unsplit solve's bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <string>
#include <vector>

#include "aux_kernels.hpp"
#include "runtime.hpp"
#include "stage_dispatch.hpp"

namespace cato {
namespace {

struct PackEntry {
  const double *prim[3][4];  // the sub-domain's primitive buffer sets
  double *send;              // 8 send regions, each in the receiving neighbour's halo-store layout
  int ni, nj, pitch, pad;
};

struct UnpackEntry {
  double *halo;              // the sub-domain's halo store
  const double *src[8];      // halo region r <- this neighbour send region (null: side not exchanged)
  int ni, nj;
};

// Region r of a send buffer goes to the neighbour in direction r (0 +x, 1 -x, 2 +y, 3 -y, 4 +x+y,
// 5 -x+y, 6 +x-y, 7 -x-y), which stores it as its halo region kOpposite[r].
constexpr int kOpposite[8] = {1, 0, 3, 2, 7, 6, 5, 4};

__device__ inline void halo_element(size_t t, int ni, int nj, int nh, int &r, int &v, int &k, int &pos) {
  r = 0;
  size_t off = t;
  while (r < 7 && off >= halo_region_size(r, ni, nj, nh)) {
    off -= halo_region_size(r, ni, nj, nh);
    ++r;
  }
  const size_t len = r < 2 ? (size_t)nj : r < 4 ? (size_t)ni : (size_t)nh;
  const size_t per_v = (size_t)nh * len;
  v = (int)(off / per_v);
  const size_t rem = off - (size_t)v * per_v;
  k = (int)(rem / len);
  pos = (int)(rem - (size_t)k * len);
}

static __global__ __launch_bounds__(256) void k_halo_pack(const PackEntry *tab, int pset, int nh) {
  const PackEntry e = tab[blockIdx.y];
  const int ni = e.ni, nj = e.nj;
  const size_t total = halo_store_size(ni, nj, nh);
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
    int r, v, k, pos;
    halo_element(t, ni, nj, nh, r, v, k, pos);
    // the cell of this sub-domain that the receiver's ghost (layer k, position pos) mirrors
    int ci, cj;
    switch (r) {
      case 0: ci = ni - nh + 1 + k; cj = pos + 1; break;            // last columns -> +x neighbour's -x strip
      case 1: ci = 1 + k; cj = pos + 1; break;                      // first columns -> -x neighbour's +x strip
      case 2: ci = pos + 1; cj = nj - nh + 1 + k; break;            // top rows -> +y neighbour's -y strip
      case 3: ci = pos + 1; cj = 1 + k; break;                      // bottom rows -> -y neighbour's +y strip
      case 4: ci = ni - nh + 1 + pos; cj = nj - nh + 1 + k; break;  // corner blocks: pos = kx, k = ky
      case 5: ci = 1 + pos; cj = nj - nh + 1 + k; break;
      case 6: ci = ni - nh + 1 + pos; cj = 1 + k; break;
      default: ci = 1 + pos; cj = 1 + k; break;
    }
    e.send[t] = e.prim[pset][v][(size_t)(cj - 1) * e.pitch + (size_t)(ci - 1)];
  }
}

static __global__ __launch_bounds__(256) void k_halo_unpack(const UnpackEntry *tab, int nh) {
  const UnpackEntry e = tab[blockIdx.y];
  const int ni = e.ni, nj = e.nj;
  const size_t total = halo_store_size(ni, nj, nh);
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
    int r, v, k, pos;
    halo_element(t, ni, nj, nh, r, v, k, pos);
    const double *src = e.src[r];
    if (src) e.halo[t] = src[t - halo_region_offset(r, ni, nj, nh)];
  }
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace
}  // namespace cato

using namespace cato;

struct cato_split {
  cato_config cfg{};
  int ni = 0, nj = 0, nh = 0, px = 0, py = 0, n = 0;
  std::vector<Solve> sub;
  std::vector<int> i0, j0;
  std::vector<std::array<int, 8>> nbr;
  std::vector<double *> halo, send;
  char *mem = nullptr;
  PackEntry *d_pack = nullptr;
  UnpackEntry *d_unpack = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double xch_ms = 0.0;
  long long xch_n = 0;
};

namespace {

#define SPCK(expr)                             \
  do {                                         \
    int _rc = hip_check((expr), #expr);        \
    if (_rc != CATO_OK) return _rc;            \
  } while (0)

// Primitive set index k of sub-domain 0 whose first field is p.
int pset_of(const cato_split *S, const double *p) {
  for (int k = 0; k < 3; ++k)
    if (S->sub[0].P[k][0] == p) return k;
  return -1;
}

// Exchange the halos of primitive set pset of every sub-domain; timed with events.
int exchange(cato_split *S, int pset) {
  if (pset < 0) return fail(CATO_ERR_INVALID, "internal: unknown primitive set");
  (void)hipEventRecord(S->ev0, S->stream);
  size_t most = 0;
  for (int s = 0; s < S->n; ++s) most = std::max(most, halo_store_size(S->sub[s].ni, S->sub[s].nj, S->nh));
  const dim3 grid((unsigned)std::max<size_t>(1, std::min<size_t>((most + 255) / 256, 64)), (unsigned)S->n);
  hipLaunchKernelGGL(k_halo_pack, grid, dim3(256), 0, S->stream, S->d_pack, pset, S->nh);
  SPCK(hipGetLastError());
  hipLaunchKernelGGL(k_halo_unpack, grid, dim3(256), 0, S->stream, S->d_unpack, S->nh);
  SPCK(hipGetLastError());
  (void)hipEventRecord(S->ev1, S->stream);
  return CATO_OK;
}

void collect_exchange_time(cato_split *S) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, S->ev0, S->ev1) == hipSuccess) {
    S->xch_ms += ms;
    S->xch_n += 1;
  }
}

int check_split(cato_split *S) { return S ? CATO_OK : fail(CATO_ERR_INVALID, "null split solve"); }

}  // namespace

extern "C" {

int cato_split_destroy(cato_split *S) {
  if (!S) return CATO_OK;
  if (S->stream) (void)hipStreamSynchronize(S->stream);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  if (S->mem) (void)hipFree(S->mem);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
  return CATO_OK;
}

int cato_split_create(const cato_config *cfg, const cato_geom *geom, int px, int py, cato_split **out) {
  if (!cfg || !geom || !out) return fail(CATO_ERR_INVALID, "null argument");
  *out = nullptr;
  if (px < 1 || py < 1 || px * py > 64) return fail(CATO_ERR_INVALID, "split: 1 <= px, py and px * py <= 64");
  if (cfg->precision != CATO_PRECISION_FP64)
    return fail(CATO_ERR_UNSUPPORTED, "split: the exchanged halos carry the FP64 form's stored primitives");
  const int bcs[4] = {cfg->bc_plus_x, cfg->bc_minus_x, cfg->bc_plus_y, cfg->bc_minus_y};
  for (int k = 0; k < 4; ++k)
    if (bcs[k] == BC_OUTFLOW || bcs[k] == BC_PRESSURE_INPUT)
      return fail(CATO_ERR_UNSUPPORTED, "split: outflow and pressure-input sides are not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CATO_ERR_NO_DEVICE, "no HIP device");
  Solve whole;
  int rc = whole.init(*cfg, *geom);  // validates the configuration and the grid as one solve
  if (rc) return rc;
  const int ni = geom->ni, nj = geom->nj, nh = cfg->n_ghost_layers;
  if (ni / px < nh || nj / py < nh) return fail(CATO_ERR_INVALID, "split: every sub-domain needs >= nh cells a side");
  cato_split *S = new cato_split();
  S->cfg = *cfg;
  S->ni = ni;
  S->nj = nj;
  S->nh = nh;
  S->px = px;
  S->py = py;
  S->n = px * py;
  S->sub.resize(S->n);
  S->i0.resize(S->n);
  S->j0.resize(S->n);
  S->nbr.resize(S->n);
  // near-equal blocks (the first ni % px columns one wider), sub-domain s = a + px * b
  std::vector<int> xs(px + 1), ys(py + 1);
  for (int a = 0; a <= px; ++a) xs[a] = a * (ni / px) + std::min(a, ni % px);
  for (int b = 0; b <= py; ++b) ys[b] = b * (nj / py) + std::min(b, nj % py);
  const int nnx = ni + 2 * nh + 1;
  const bool per_x = cfg->bc_plus_x == BC_PERIODIC, per_y = cfg->bc_plus_y == BC_PERIODIC;
  auto step = [](int a, int d, int np, bool periodic) -> int {
    int t = a + d;
    if (t >= 0 && t < np) return t;
    if (!periodic || np == 1) return -1;
    return (t + np) % np;
  };
  size_t bytes = 0;
  std::vector<size_t> off(S->n);
  for (int b = 0; b < py; ++b)
    for (int a = 0; a < px; ++a) {
      const int s = a + px * b;
      const int sni = xs[a + 1] - xs[a], snj = ys[b + 1] - ys[b];
      S->i0[s] = xs[a] + 1;
      S->j0[s] = ys[b] + 1;
      // the sub-domain's nodes with nh ghost nodes: a window of the global node arrays
      const int mx = sni + 2 * nh + 1, my = snj + 2 * nh + 1;
      std::vector<double> nx((size_t)mx * my), ny((size_t)mx * my);
      for (int q = 0; q < my; ++q)
        for (int p = 0; p < mx; ++p) {
          const size_t g = (size_t)(xs[a] + p) + (size_t)(ys[b] + q) * nnx;
          nx[(size_t)p + (size_t)q * mx] = geom->node_x[g];
          ny[(size_t)p + (size_t)q * mx] = geom->node_y[g];
        }
      cato_geom sg{sni, snj, nx.data(), ny.data()};
      if ((rc = S->sub[s].init(*cfg, sg))) {
        delete S;
        return rc;
      }
      const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
      for (int r = 0; r < 8; ++r) {
        const int ta = step(a, dirs[r][0], px, per_x), tb = step(b, dirs[r][1], py, per_y);
        const bool need_x = dirs[r][0] != 0, need_y = dirs[r][1] != 0;
        S->nbr[s][r] = ((need_x && ta < 0) || (need_y && tb < 0)) ? -1 : (need_x ? ta : a) + px * (need_y ? tb : b);
      }
      int mask = 0;
      for (int r = 0; r < 4; ++r)
        if (S->nbr[s][r] >= 0) mask |= 1 << r;
      S->sub[s].halo_mask = mask;
      off[s] = bytes;
      bytes += align256(S->sub[s].device_bytes()) + 2 * align256(halo_store_size(sni, snj, nh) * sizeof(double));
    }
  const size_t tab_off = bytes;
  bytes += align256(S->n * sizeof(PackEntry)) + align256(S->n * sizeof(UnpackEntry));
  if ((rc = hip_check(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking), "split stream")) ||
      (rc = hip_check(hipMalloc((void **)&S->mem, bytes), "split device allocation")) ||
      (rc = hip_check(hipEventCreate(&S->ev0), "split event")) || (rc = hip_check(hipEventCreate(&S->ev1), "split event"))) {
    cato_split_destroy(S);
    return rc;
  }
  S->halo.resize(S->n);
  S->send.resize(S->n);
  for (int s = 0; s < S->n; ++s) {
    Solve &v = S->sub[s];
    char *base = S->mem + off[s];
    v.bind(base);
    const size_t hs = align256(halo_store_size(v.ni, v.nj, nh) * sizeof(double));
    S->halo[s] = reinterpret_cast<double *>(base + align256(v.device_bytes()));
    S->send[s] = reinterpret_cast<double *>(base + align256(v.device_bytes()) + hs);
    v.halo = v.halo_mask ? S->halo[s] : nullptr;
    if ((rc = v.upload_geometry(S->stream))) {
      cato_split_destroy(S);
      return rc;
    }
  }
  std::vector<PackEntry> pk(S->n);
  std::vector<UnpackEntry> up(S->n);
  for (int s = 0; s < S->n; ++s) {
    const Solve &v = S->sub[s];
    for (int k = 0; k < 3; ++k)
      for (int f = 0; f < 4; ++f) pk[s].prim[k][f] = v.P[k][f];
    pk[s].send = S->send[s];
    pk[s].ni = v.ni;
    pk[s].nj = v.nj;
    pk[s].pitch = v.pitch;
    up[s].halo = S->halo[s];
    up[s].ni = v.ni;
    up[s].nj = v.nj;
    for (int r = 0; r < 8; ++r) {
      const int q = S->nbr[s][r];
      // a corner is read only where both of its sides are exchanged
      const bool used = r < 4 || ((v.halo_mask & (r == 4 || r == 6 ? 1 : 2)) && (v.halo_mask & (r < 6 ? 4 : 8)));
      up[s].src[r] = (q >= 0 && used)
                         ? S->send[q] + halo_region_offset(kOpposite[r], S->sub[q].ni, S->sub[q].nj, nh)
                         : nullptr;
    }
  }
  S->d_pack = reinterpret_cast<PackEntry *>(S->mem + tab_off);
  S->d_unpack = reinterpret_cast<UnpackEntry *>(S->mem + tab_off + align256(S->n * sizeof(PackEntry)));
  if ((rc = hip_check(hipMemcpyAsync(S->d_pack, pk.data(), pk.size() * sizeof(PackEntry), hipMemcpyHostToDevice,
                                     S->stream), "split tables")) ||
      (rc = hip_check(hipMemcpyAsync(S->d_unpack, up.data(), up.size() * sizeof(UnpackEntry), hipMemcpyHostToDevice,
                                     S->stream), "split tables")) ||
      (rc = hip_check(hipStreamSynchronize(S->stream), "split create"))) {
    cato_split_destroy(S);
    return rc;
  }
  *out = S;
  return CATO_OK;
}

int cato_split_set_primitive(cato_split *S, const double *rho, const double *u, const double *v, const double *p,
                             int ld, int halo) {
  int rc = check_split(S);
  if (rc) return rc;
  if (!rho || !u || !v || !p) return fail(CATO_ERR_INVALID, "null field pointer");
  if (halo < 0 || ld < S->ni + 2 * halo) return fail(CATO_ERR_INVALID, "bad leading dimension / halo");
  const double *H[4] = {rho, u, v, p};
  for (int s = 0; s < S->n; ++s) {
    Solve &w = S->sub[s];
    const size_t o = (size_t)(S->i0[s] - 1) + (size_t)(S->j0[s] - 1) * (size_t)ld;
    for (int f = 0; f < 4; ++f)
      if ((rc = field_up(S->stream, w, w.P[0][f], H[f] + o, ld, halo, nullptr))) return rc;
    hipLaunchKernelGGL(k_prim_to_cons, dim3((unsigned)((w.ni + 255) / 256), (unsigned)w.nj), dim3(256), 0,
                       S->stream, w.fgeom(), w.P[0][0], w.P[0][1], w.P[0][2], w.P[0][3], w.U[w.cur][0],
                       w.U[w.cur][1], w.U[w.cur][2], w.U[w.cur][3]);
    SPCK(hipGetLastError());
    w.prims_valid = true;
  }
  if ((rc = exchange(S, 0))) return rc;
  return hip_check(hipStreamSynchronize(S->stream), "split set_primitive");
}

int cato_split_get_conserved(cato_split *S, double *rho, double *rho_u, double *rho_v, double *rho_E, int ld,
                             int halo) {
  int rc = check_split(S);
  if (rc) return rc;
  if (halo < 0 || ld < S->ni + 2 * halo) return fail(CATO_ERR_INVALID, "bad leading dimension / halo");
  double *H[4] = {rho, rho_u, rho_v, rho_E};
  for (int s = 0; s < S->n; ++s) {
    Solve &w = S->sub[s];
    const size_t o = (size_t)(S->i0[s] - 1) + (size_t)(S->j0[s] - 1) * (size_t)ld;
    for (int f = 0; f < 4; ++f)
      if (H[f] && (rc = field_down(S->stream, w, H[f] + o, w.U[w.cur][f], ld, halo, nullptr))) return rc;
  }
  return hip_check(hipStreamSynchronize(S->stream), "split get_conserved");
}

int cato_split_compute_dt(cato_split *S, double *dt) {
  int rc = check_split(S);
  if (rc) return rc;
  std::vector<double> r((size_t)2 * S->n);
  for (int s = 0; s < S->n; ++s) {
    Solve &w = S->sub[s];
    const int nb = w.tiles_x * w.tiles_y;
    hipLaunchKernelGGL(k_cfl_partial, dim3((unsigned)nb), dim3(kThreads), 0, S->stream, w.fgeom(), 1, w.P[0][0],
                       w.P[0][1], w.P[0][2], w.P[0][3], w.block_min, w.block_err);
    hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(1024), 0, S->stream, nb, w.block_min, w.block_err, w.result,
                       (LoopState *)nullptr, 0, ConvOut{});
    SPCK(hipGetLastError());
    SPCK(hipMemcpyAsync(&r[(size_t)2 * s], w.result, 2 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
  }
  SPCK(hipStreamSynchronize(S->stream));
  double m = r[0];
  for (int s = 1; s < S->n; ++s) m = (r[(size_t)2 * s] < m) ? r[(size_t)2 * s] : m;
  if (dt) *dt = m;
  return CATO_OK;
}

int cato_split_step(cato_split *S, double dt, int *error_code) {
  int rc = check_split(S);
  if (rc) return rc;
  if (!(dt >= 2.2250738585072014e-308)) return fail(CATO_ERR_INVALID, "The timestep dt is < tiny(1.0_rk)");
  std::vector<std::vector<StageArgs>> st(S->n);
  for (int s = 0; s < S->n; ++s) st[s] = S->sub[s].step_stages(nullptr, dt);
  const size_t ns = st[0].size();
  for (size_t m = 0; m < ns; ++m) {
    for (int s = 0; s < S->n; ++s) {
      const StageArgs &a = st[s][m];
      if ((rc = validate_stage_args(a))) return rc;
      S->sub[s].launcher(dim3((unsigned)(S->sub[s].tiles_x * S->sub[s].tiles_y)), S->stream, a);
      SPCK(hipGetLastError());
    }
    // the next stage (or the next step) reads these primitives, halos included
    if ((rc = exchange(S, pset_of(S, st[0][m].prim_out[0])))) return rc;
    SPCK(hipStreamSynchronize(S->stream));
    collect_exchange_time(S);
  }
  std::vector<double> r((size_t)2 * S->n);
  for (int s = 0; s < S->n; ++s) {
    Solve &w = S->sub[s];
    hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(1024), 0, S->stream, w.tiles_x * w.tiles_y, w.block_min,
                       w.block_err, w.result, (LoopState *)nullptr, 0, ConvOut{});
    SPCK(hipGetLastError());
    SPCK(hipMemcpyAsync(&r[(size_t)2 * s], w.result, 2 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    w.prims_valid = false;
  }
  SPCK(hipStreamSynchronize(S->stream));
  int f = 0;
  for (int s = 0; s < S->n; ++s) f |= (int)r[(size_t)2 * s + 1];
  if (error_code) *error_code = (f & 1) ? 3 : (f & 2) ? 1 : (f & 4) ? 2 : 0;
  return CATO_OK;
}

int cato_split_exchange_time(cato_split *S, double *ms_total, long long *exchanges) {
  int rc = check_split(S);
  if (rc) return rc;
  if (ms_total) *ms_total = S->xch_ms;
  if (exchanges) *exchanges = S->xch_n;
  return CATO_OK;
}

}  // extern "C"
