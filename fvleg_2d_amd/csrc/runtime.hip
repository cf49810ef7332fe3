// Host runtime of the CATO HIP library: geometry, device-resident state,
// the stage sequences of the SSP/RK integrators, the device-side time loop,
// operator-level entry points and the C ABI of include/cato_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "aux_kernels.hpp"
#include "runtime.hpp"
#include "stage_dispatch.hpp"

namespace cato {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return CATO_OK;
  return fail(CATO_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCK(expr)                                     \
  do {                                                  \
    int _rc = hip_check((expr), #expr);                 \
    if (_rc != CATO_OK) return _rc;                     \
  } while (0)

// ---------------------------------------------------------------------------
// Device memory of the allocation-free entry point (cato_compute_dt_noalloc):
// the per-workgroup CFL partial minima and sanity flags and the reduced
// result live in this fixed array instead of an allocation.  Every GPU that
// loads the library holds its own copy (about 48 KiB) until the process ends;
// host code reaches it only through the per-device address that
// hipGetSymbolAddress returns (noalloc_scratch), never through the symbol.
// ---------------------------------------------------------------------------
constexpr int kNoAllocBlocks = 4096;  // partials per call: larger grids stream through 4096 workgroups
constexpr size_t kNoAllocMinOff = 0;
constexpr size_t kNoAllocErrOff = kNoAllocMinOff + kNoAllocBlocks * sizeof(double);
constexpr size_t kNoAllocResOff = kNoAllocErrOff + kNoAllocBlocks * sizeof(int);
constexpr size_t kNoAllocBytes = kNoAllocResOff + 2 * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current: the device address of `scratch`.
static int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(CATO_ERR_INVALID, "device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return CATO_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hip_check(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
  g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
  return CATO_OK;
}

static unsigned char *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ---------------------------------------------------------------------------
// Geometry: quad_cell.f90:69-229 per cell.  A rectilinear block (node_x
// depends on i only, node_y on j only, axis-aligned normals) keeps four 1D
// arrays from which the fused kernel reproduces every cell's geometry exactly;
// any other grid keeps the full per-cell geometry (Geometry1D::q).
// ---------------------------------------------------------------------------
static inline double hmax2(double a, double b) { return (b > a) ? b : a; }
static inline double hmin2(double a, double b) { return (b < a) ? b : a; }

// quad_cell%initialize for every halo'd cell: round-off snapping of the node
// coordinates (quad_cell.f90:78-91), volume (160-175), edge lengths with their
// equalisation (100-112), min_dx / min_dy (119-123) and the unit normals of the
// right and top edges from the edge midpoints (193-229, vector.f90:58-96).
static int build_quad_geometry(int ni, int nj, int nh, const double *nx, const double *ny, Geometry1D &g) {
  const double EPS = 2.220446049250313080847e-16;
  const int nnx = ni + 2 * nh + 1;
  const int ncx = ni + 2 * nh, ncy = nj + 2 * nh;
  const size_t n = (size_t)ncx * ncy;
  auto NX = [&](int a, int b) { return nx[(size_t)a + (size_t)b * nnx]; };
  auto NY = [&](int a, int b) { return ny[(size_t)a + (size_t)b * nnx]; };
  g.quad = true;
  g.q.assign((size_t)Q_COUNT * n, 0.0);
  g.dxi.assign(ncx, 0.0);
  g.lxi.assign(ncx, 0.0);
  g.dyj.assign(ncy, 0.0);
  g.lyj.assign(ncy, 0.0);
  for (int b = 0; b < ncy; ++b)
    for (int a = 0; a < ncx; ++a) {
      double x[4] = {NX(a, b), NX(a + 1, b), NX(a + 1, b + 1), NX(a, b + 1)};
      double y[4] = {NY(a, b), NY(a + 1, b), NY(a + 1, b + 1), NY(a, b + 1)};
      if (std::fabs(y[1] - y[0]) < EPS) y[1] = y[0];
      if (std::fabs(y[3] - y[2]) < EPS) y[3] = y[2];
      if (std::fabs(x[2] - x[1]) < EPS) x[2] = x[1];
      if (std::fabs(x[3] - x[0]) < EPS) x[3] = x[0];
      const double vol = (x[1] - x[0]) * (y[3] - y[0]) - (x[3] - x[0]) * (y[1] - y[0]);
      if (!(vol > 0.0)) return fail(CATO_ERR_INVALID, "non-positive cell volume (nodes must run counter-clockwise)");
      double el[4];
      for (int e = 0; e < 4; ++e) {
        const int h = (e + 1) & 3;
        const double ex = x[h] - x[e], ey = y[h] - y[e];
        el[e] = std::sqrt(ex * ex + ey * ey);
      }
      double mx = el[0], mn = el[0];
      for (int e = 1; e < 4; ++e) {
        mx = hmax2(mx, el[e]);
        mn = hmin2(mn, el[e]);
      }
      if (mx - mn < 2.0 * EPS) el[0] = el[1] = el[2] = el[3] = mx;
      const size_t c = (size_t)a + (size_t)b * ncx;
      for (int e = 0; e < 4; ++e) g.q[(size_t)(Q_LEN0 + e) * n + c] = el[e];
      g.q[(size_t)Q_VOL * n + c] = vol;
      g.q[(size_t)Q_MDX * n + c] = hmin2(std::fabs(x[1] - x[0]), std::fabs(x[2] - x[3]));
      g.q[(size_t)Q_MDY * n + c] = hmin2(std::fabs(y[3] - y[0]), std::fabs(y[2] - y[1]));
      for (int e = 1; e <= 2; ++e) {  // right edge (i-flux) and top edge (j-flux)
        const int h = e + 1;
        const double mxp = (x[h] + x[e]) / 2.0, myp = (y[h] + y[e]) / 2.0;
        const double ddx = x[h] - mxp, ddy = y[h] - myp;
        const double vx = (mxp + ddy) - mxp, vy = (myp - ddx) - myp;
        const double len = std::sqrt(vx * vx + vy * vy);
        g.q[(size_t)(e == 1 ? Q_NRX : Q_NTX) * n + c] = vx / len;
        g.q[(size_t)(e == 1 ? Q_NRY : Q_NTY) * n + c] = vy / len;
      }
    }
  return CATO_OK;
}

// The stage epilogue's cell geometry (stage_kernel.hpp, phase 3) when every
// interior cell has the same dxi, lxi, dyj and lyj, in the kernel's operations
// (IEEE double on both sides, so the same bits): edge-length equalisation
// (quad_cell.f90:110-112) over (l1, l2, l1, l2), volume dx dy, and the exact
// reciprocal of a power-of-two volume.
static void uniform_cell_geometry(int ni, int nj, int nh, Geometry1D &g) {
  g.uniform = false;
  for (int a = nh; a < nh + ni; ++a)
    if (g.dxi[a] != g.dxi[nh] || g.lxi[a] != g.lxi[nh]) return;
  for (int b = nh; b < nh + nj; ++b)
    if (g.dyj[b] != g.dyj[nh] || g.lyj[b] != g.lyj[nh]) return;
  const double dx = g.dxi[nh], dy = g.dyj[nh], l1 = g.lxi[nh], l2 = g.lyj[nh];
  const double mx = (l2 > l1) ? l2 : l1, mn = (l2 < l1) ? l2 : l1;
  double lb = l1, lr = l2, lt = l1, ll = l2;
  if (mx - mn < 2.0 * 2.220446049250313080847e-16) lb = lr = lt = ll = mx;
  const double vol = dx * dy;
  uint64_t vb;
  std::memcpy(&vb, &vol, 8);
  const uint64_t ve = (vb >> 52) & 0x7FF;
  const bool pow2 = (vb & 0x000FFFFFFFFFFFFFull) == 0 && ve > 1 && ve < 2046;
  const uint64_t ib = (vb & 0x8000000000000000ull) | ((2046 - ve) << 52);
  double inv;
  std::memcpy(&inv, &ib, 8);
  const double u[8] = {dx, dy, lb, lr, lt, ll, vol, pow2 ? inv : 0.0};
  for (int k = 0; k < 8; ++k) g.u[k] = u[k];
  g.u_pow2 = pow2;
  g.uniform = true;
}

int build_geometry(int ni, int nj, int nh, const double *nx, const double *ny, Geometry1D &g) {
  const int nnx = ni + 2 * nh + 1, nny = nj + 2 * nh + 1;
  auto NX = [&](int a, int b) { return nx[(size_t)a + (size_t)b * nnx]; };
  auto NY = [&](int a, int b) { return ny[(size_t)a + (size_t)b * nnx]; };
  g.quad = false;
  g.q.clear();
  for (int b = 0; b < nny; ++b)
    for (int a = 0; a < nnx; ++a) {
      if (NX(a, b) != NX(a, 0) || NY(a, b) != NY(0, b)) return build_quad_geometry(ni, nj, nh, nx, ny, g);
    }
  const int ncx = ni + 2 * nh, ncy = nj + 2 * nh;
  g.dxi.resize(ncx);
  g.lxi.resize(ncx);
  g.dyj.resize(ncy);
  g.lyj.resize(ncy);
  for (int a = 0; a < ncx; ++a) {
    double d = NX(a + 1, 0) - NX(a, 0);
    g.dxi[a] = d;
    g.lxi[a] = std::sqrt(d * d + 0.0 * 0.0);
    if (!(d > 0.0)) return fail(CATO_ERR_INVALID, "non-positive cell width in x");
  }
  for (int b = 0; b < ncy; ++b) {
    double d = NY(0, b + 1) - NY(0, b);
    g.dyj[b] = d;
    g.lyj[b] = std::sqrt(0.0 * 0.0 + d * d);
    if (!(d > 0.0)) return fail(CATO_ERR_INVALID, "non-positive cell width in y");
  }
  // Cross-check the separable form against the full quad-cell algorithm on
  // every cell; a grid whose volumes or normals it would not reproduce exactly
  // takes the per-cell geometry instead.
  for (int b = 0; b < ncy; ++b)
    for (int a = 0; a < ncx; ++a) {
      double x[4] = {NX(a, b), NX(a + 1, b), NX(a + 1, b + 1), NX(a, b + 1)};
      double y[4] = {NY(a, b), NY(a + 1, b), NY(a + 1, b + 1), NY(a, b + 1)};
      double vol = (x[1] - x[0]) * (y[3] - y[0]) - (x[3] - x[0]) * (y[1] - y[0]);
      if (vol != g.dxi[a] * g.dyj[b]) return build_quad_geometry(ni, nj, nh, nx, ny, g);
      const int head[4] = {1, 2, 3, 0};
      for (int e = 1; e <= 2; ++e) {  // RIGHT and TOP edges carry the i and j fluxes
        double mx = (x[head[e]] + x[e]) / 2.0, my = (y[head[e]] + y[e]) / 2.0;
        double ddx = x[head[e]] - mx, ddy = y[head[e]] - my;
        double vx = (mx + ddy) - mx, vy = (my - ddx) - my;
        double len = std::sqrt(vx * vx + vy * vy);
        double ex = (e == 1) ? 1.0 : 0.0, ey = (e == 1) ? 0.0 : 1.0;
        if (vx / len != ex || vy / len != ey) return build_quad_geometry(ni, nj, nh, nx, ny, g);
      }
    }
  uniform_cell_geometry(ni, nj, nh, g);
  return CATO_OK;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
static int validate_config(const cato_config &c) {
  if (c.flux < 0 || c.flux >= FLUX_COUNT) return fail(CATO_ERR_INVALID, "unknown flux solver");
  if (c.limiter < 0 || c.limiter >= LIM_COUNT) return fail(CATO_ERR_INVALID, "unknown limiter");
  if (c.integrator < CATO_RK_SSP2 || c.integrator > CATO_RK_2_5)
    return fail(CATO_ERR_INVALID, "unknown time integration scheme (ssp_rk43 has no body in the reference)");
  if (c.precision != CATO_PRECISION_FP64 && c.precision != CATO_PRECISION_FP32_FLUX)
    return fail(CATO_ERR_INVALID, "unknown precision");
  int need = (c.limiter == LIM_MLP5) ? 3 : 2;
  if (c.n_ghost_layers != need)
    return fail(CATO_ERR_INVALID,
                "The number of required ghost cell layers doesn't match the edge interpolation order");
  const int bcs[4] = {c.bc_plus_x, c.bc_minus_x, c.bc_plus_y, c.bc_minus_y};
  bool outflow = false;
  for (int k = 0; k < 4; ++k) {
    if (bcs[k] < 0 || bcs[k] > 4) return fail(CATO_ERR_INVALID, "unsupported boundary condition type");
    outflow = outflow || bcs[k] == BC_OUTFLOW;
    if (k > 0 && bcs[k] == BC_PRESSURE_INPUT)  // pressure_input_bc.f90:377-381
      return fail(CATO_ERR_INVALID, "Unsupported pressure_bc location (pressure_input is a +x boundary)");
  }
  if (outflow && !(c.bc_pressure > 0.0))
    return fail(CATO_ERR_INVALID, "Outlet ambient pressure is <= 0!");  // outlet_bc.f90:57-62
  // A y outlet shares the pressure input's priority and is applied after it, over
  // rows the pressure input has already read (its column maximum would see them).
  if (c.bc_plus_x == BC_PRESSURE_INPUT && (c.bc_plus_y == BC_OUTFLOW || c.bc_minus_y == BC_OUTFLOW))
    return fail(CATO_ERR_UNSUPPORTED, "pressure_input at +x with an outlet at +y or -y");
  if ((c.bc_plus_x == BC_PERIODIC) != (c.bc_minus_x == BC_PERIODIC) ||
      (c.bc_plus_y == BC_PERIODIC) != (c.bc_minus_y == BC_PERIODIC))
    return fail(CATO_ERR_UNSUPPORTED, "a periodic boundary must be periodic on both sides of its axis");
  if (!(c.gamma > 1.0)) return fail(CATO_ERR_INVALID, "polytropic_index must be > 1");
  if (!(c.cfl > 0.0)) return fail(CATO_ERR_INVALID, "cfl must be > 0");
  if (c.tile >= kNumTiles) return fail(CATO_ERR_INVALID, "unknown tile id");
  return CATO_OK;
}

int default_tile(const cato_config &c, int ni, int nj) {
  if (c.tile >= 0) return c.tile;
  (void)ni;
  (void)nj;
  return 0;  // 32 x 8
}

int Solve::init(const cato_config &c, const cato_geom &g, bool allow_quad) {
  int rc = validate_config(c);
  if (rc) return rc;
  if (g.ni < c.n_ghost_layers || g.nj < c.n_ghost_layers)
    return fail(CATO_ERR_INVALID, "grid must have at least n_ghost_layers interior cells per direction");
  if (!g.node_x || !g.node_y) return fail(CATO_ERR_INVALID, "null node coordinates");
  cfg = c;
  ni = g.ni;
  nj = g.nj;
  nh = c.n_ghost_layers;
  // a +x outflow side keeps its ghost column (ni+1) in the primitive arrays
  pitch = (ni + (ghost_column_bc(c.bc_plus_x) ? 1 : 0) + 15) & ~15;
  rc = build_geometry(ni, nj, nh, g.node_x, g.node_y, geo);
  if (rc) return rc;
  if (geo.quad && !allow_quad)
    return fail(CATO_ERR_UNSUPPORTED,
                "general quadrilateral grids run as single solves (cato_ctx); this form takes rectilinear grids");
  // quad cells: the one tile shape their kernels are instantiated for
  tile = geo.quad ? kQuadTile : default_tile(c, ni, nj);
  tiles_x = (ni + kTiles[tile].tx - 1) / kTiles[tile].tx;
  tiles_y = (nj + kTiles[tile].ty - 1) / kTiles[tile].ty;
  launcher = stage_launcher(c.flux, c.limiter, c.precision, tile, geo.quad ? 1 : 0);
  if (!launcher) return fail(CATO_ERR_INVALID, "no stage kernel for this configuration");
  return CATO_OK;
}

size_t Solve::field_elems() const { return (size_t)pitch * nj; }

size_t Solve::device_bytes() const {
  size_t f = field_elems();
  size_t nb = (size_t)tiles_x * tiles_y;
  size_t geo_n = (size_t)2 * (ni + 2 * nh) + 2 * (nj + 2 * nh);
  return (12 + 12) * f * sizeof(double) + geo_n * sizeof(double) + (size_t)Q_COUNT * quad_elems() * sizeof(double) +
         nb * (sizeof(double) + sizeof(int)) + sizeof(LoopState) + 4 * sizeof(double) + 1024;
}

// Carve the device arrays of one solve out of `base` (device_bytes() long).
void Solve::bind(char *base) {
  size_t f = field_elems();
  double *p = reinterpret_cast<double *>(base);
  for (int s = 0; s < 3; ++s)
    for (int v = 0; v < 4; ++v) {
      U[s][v] = p;
      p += f;
    }
  for (int s = 0; s < 3; ++s)
    for (int v = 0; v < 4; ++v) {
      P[s][v] = p;
      p += f;
    }
  dxi = p; p += ni + 2 * nh;
  lxi = p; p += ni + 2 * nh;
  dyj = p; p += nj + 2 * nh;
  lyj = p; p += nj + 2 * nh;
  quad = geo.quad ? p : nullptr;
  p += (size_t)Q_COUNT * quad_elems();
  block_min = p; p += (size_t)tiles_x * tiles_y;
  result = p; p += 4;
  char *c = reinterpret_cast<char *>(p);
  c = reinterpret_cast<char *>(((uintptr_t)c + 255) & ~(uintptr_t)255);
  loop = reinterpret_cast<LoopState *>(c);
  c += sizeof(LoopState);
  block_err = reinterpret_cast<int *>(c);
}

int Solve::upload_geometry(hipStream_t s) {
  HIPCK(hipMemcpyAsync(dxi, geo.dxi.data(), geo.dxi.size() * 8, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(lxi, geo.lxi.data(), geo.lxi.size() * 8, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dyj, geo.dyj.data(), geo.dyj.size() * 8, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(lyj, geo.lyj.data(), geo.lyj.size() * 8, hipMemcpyHostToDevice, s));
  if (geo.quad) HIPCK(hipMemcpyAsync(quad, geo.q.data(), geo.q.size() * 8, hipMemcpyHostToDevice, s));
  HIPCK(hipMemsetAsync(block_err, 0, (size_t)tiles_x * tiles_y * sizeof(int), s));
  LoopState L{};
  HIPCK(hipMemcpyAsync(loop, &L, sizeof(L), hipMemcpyHostToDevice, s));
  return CATO_OK;
}

FieldGeom Solve::fgeom() const {
  FieldGeom g;
  g.ni = ni;
  g.nj = nj;
  g.nh = nh;
  g.pitch = pitch;
  g.bc[0] = cfg.bc_plus_x;
  g.bc[1] = cfg.bc_minus_x;
  g.bc[2] = cfg.bc_plus_y;
  g.bc[3] = cfg.bc_minus_y;
  g.gamma = cfg.gamma;
  gas_constants(cfg.gamma, g.gas);
  g.cfl = cfg.cfl;
  g.bc_pressure = cfg.bc_pressure;
  g.dxi = dxi;
  g.dyj = dyj;
  g.quad = quad;
  g.quad_ncx = ni + 2 * nh;
  g.quad_n = (long long)quad_elems();
  return g;
}

StageArgs Solve::stage_args(int mode, double a, double b, double c, int in_buf, int u0_buf, int out_buf, int p_in,
                            int p_out, bool last, const LoopState *lp, double dt_value, bool check_input) const {
  StageArgs s{};
  s.ni = ni;
  s.nj = nj;
  s.nh = nh;
  s.pitch = pitch;
  s.mode = mode;
  s.check_input = check_input ? 1 : 0;
  s.last_stage = last ? 1 : 0;
  s.low_mach_fix = cfg.apply_low_mach_fix;
  s.fast_ops = (cfg.gamma > 1.0 && cfg.gamma <= 4.0) ? 1 : 0;  // the range bounds of fast_ops_cell assume it
  s.bc[0] = cfg.bc_plus_x;
  s.bc[1] = cfg.bc_minus_x;
  s.bc[2] = cfg.bc_plus_y;
  s.bc[3] = cfg.bc_minus_y;
  s.tiles_x = tiles_x;
  s.tiles_y = tiles_y;
  s.gamma = cfg.gamma;
  gas_constants(cfg.gamma, s.gas);
  s.cfl = cfg.cfl;
  s.bc_pressure = cfg.bc_pressure;
  s.a = a;
  s.b = b;
  s.c = c;
  for (int v = 0; v < 4; ++v) {
    s.uk[v] = in_buf >= 0 ? U[in_buf][v] : nullptr;
    s.u0[v] = u0_buf >= 0 ? U[u0_buf][v] : nullptr;
    s.prim[v] = P[p_in][v];
    s.out[v] = U[out_buf][v];
    s.prim_out[v] = p_out >= 0 ? P[p_out][v] : nullptr;
  }
  s.src_E = mode == MODE_RESIDUAL ? nullptr : src_E;
  s.dxi = dxi;
  s.lxi = lxi;
  s.dyj = dyj;
  s.lyj = lyj;
  s.ugeom = geo.uniform ? 1 : 0;
  s.ugeom_pow2 = geo.u_pow2 ? 1 : 0;
  for (int k = 0; k < 8; ++k) s.ug[k] = geo.u[k];
  s.loop = lp;
  s.dt_value = dt_value;
  s.block_min = block_min;
  s.block_err = block_err;
  s.halo_mask = halo_mask;
  s.halo = halo;
  s.quad = quad;
  s.quad_ncx = ni + 2 * nh;
  s.quad_n = (long long)quad_elems();
  s.fstride = (long long)field_elems();
  return s;
}

// Primitive buffer sets read and written by stage m of an ns-stage step.  P[0]
// holds the primitives of U^n at every step boundary, so the arguments of a
// step never change (one captured graph serves every step): even ns cycles
// 0->1->0->...->0, odd ns cycles 0->1->2->1->2->...->0.
static void prim_cycle(int m, int ns, int &pin, int &pout) {
  if (ns % 2 == 0) {
    pin = m % 2;
    pout = (m + 1) % 2;
    return;
  }
  pin = (m == 0) ? 0 : 1 + (m - 1) % 2;
  pout = (m == ns - 1) ? 0 : 1 + m % 2;
}

// Stage sequence of one step: ssp_rk_2_2 (mod_fluid.f90:1115-1129),
// ssp_rk_3_3 (773-802) and rk_multistage (957-966).  State buffers: cur holds
// U^n, the two others hold intermediate stages; the last stage overwrites U^n
// in place (it reads U^n only at its own cell).
std::vector<StageArgs> Solve::step_stages(const LoopState *lp, double dt_value, bool keep_un) const {
  std::vector<StageArgs> st;
  const int n = cur, A = (cur + 1) % 3, B = (cur + 2) % 3;
  // The first stage after cato_set_primitive reads the caller's primitives
  // (fluid_t%initialize keeps the deck's u, v, p) and checks them for NaN.
  const bool chk = prims_valid;
  int pi, po;
  switch (cfg.integrator) {
    case CATO_RK_SSP2:
      prim_cycle(0, 2, pi, po);
      st.push_back(stage_args(MODE_FIRST, 0, 0, 1.0, n, -1, A, pi, po, false, lp, dt_value, chk));
      prim_cycle(1, 2, pi, po);
      // keep_un: U^(n+1) into B and its primitives into P[2], both unused by the step
      st.push_back(stage_args(MODE_COMBINE, 0.5, 0.5, 0.5, A, n, keep_un ? B : n, pi, keep_un ? 2 : po, true, lp,
                              dt_value, false));
      break;
    case CATO_RK_SSP3:
      prim_cycle(0, 3, pi, po);
      st.push_back(stage_args(MODE_FIRST, 0, 0, 1.0, n, -1, A, pi, po, false, lp, dt_value, chk));
      prim_cycle(1, 3, pi, po);
      st.push_back(
          stage_args(MODE_COMBINE, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, A, n, B, pi, po, false, lp, dt_value, false));
      prim_cycle(2, 3, pi, po);
      // keep_un: U^(n+1) into A (U_1 there is read only at the cell, before the write, for the
      // convergence norms) and its primitives into P[1], which stage 2 has read
      st.push_back(stage_args(MODE_COMBINE, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, B, n, keep_un ? A : n, pi,
                              keep_un ? 1 : po, true, lp, dt_value, false));
      break;
    default: {
      static const double a13[3] = {0.1481, 0.4, 1.0}, a14[4] = {0.0833, 0.2069, 0.4265, 1.0},
                          a15[5] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0}, a23[3] = {0.1918, 0.4929, 1.0},
                          a24[4] = {0.1084, 0.2602, 0.5052, 1.0}, a25[5] = {0.0695, 0.1602, 0.2898, 0.5060, 1.0};
      const double *al;
      int ns;
      switch (cfg.integrator) {
        case CATO_RK_1_3: al = a13; ns = 3; break;
        case CATO_RK_1_4: al = a14; ns = 4; break;
        case CATO_RK_1_5: al = a15; ns = 5; break;
        case CATO_RK_2_3: al = a23; ns = 3; break;
        case CATO_RK_2_4: al = a24; ns = 4; break;
        default: al = a25; ns = 5; break;
      }
      int prev = n;
      for (int m = 0; m < ns; ++m) {
        bool last = (m == ns - 1);
        int out = last ? n : ((m % 2 == 0) ? A : B);
        prim_cycle(m, ns, pi, po);
        st.push_back(stage_args(MODE_BASE, 0, 0, al[m], prev, n, out, pi, po, last, lp, dt_value, chk && m == 0));
        prev = out;
      }
    }
  }
  if (conv_sums) {
    // calculate_convergence compares the last stage with U_1 (ssp_rk_2_2:1132,
    // ssp_rk_3_3:805; still in buffer A) or with U_0 (rk_multistage:970; read
    // at the cell before the last stage overwrites it)
    const int first = (cfg.integrator == CATO_RK_SSP2 || cfg.integrator == CATO_RK_SSP3) ? A : n;
    StageArgs &l = st.back();
    for (int v = 0; v < 4; ++v) l.conv_first[v] = U[first][v];
    l.block_conv = conv_sums;
  }
  return st;
}

int field_up(hipStream_t st, const Solve &s, double *dev, const double *host, int ld, int halo, double *staging) {
  if (!host) return fail(CATO_ERR_INVALID, "null field pointer");
  if (halo < 0 || ld < s.ni + 2 * halo) return fail(CATO_ERR_INVALID, "bad leading dimension / halo");
  const size_t n = (size_t)s.ni * s.nj;
  if (ld == s.ni && halo == 0) {
    if (s.pitch == s.ni) return hip_check(hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, st), "upload");
    if (staging) {
      int rc = hip_check(hipMemcpyAsync(staging, host, n * 8, hipMemcpyHostToDevice, st), "upload");
      if (rc) return rc;
      hipLaunchKernelGGL(k_repack, dim3((unsigned)((s.ni + 255) / 256), (unsigned)s.nj), dim3(256), 0, st,
                         (const double *)staging, s.ni, dev, s.pitch, s.ni);
      return hip_check(hipGetLastError(), "repack");
    }
  }
  const double *p = host + (size_t)halo + (size_t)halo * ld;
  return hip_check(hipMemcpy2DAsync(dev, (size_t)s.pitch * 8, p, (size_t)ld * 8, (size_t)s.ni * 8, (size_t)s.nj,
                                    hipMemcpyHostToDevice, st),
                   "upload");
}

int field_down(hipStream_t st, const Solve &s, double *host, const double *dev, int ld, int halo, double *staging) {
  if (!host) return fail(CATO_ERR_INVALID, "null field pointer");
  if (halo < 0 || ld < s.ni + 2 * halo) return fail(CATO_ERR_INVALID, "bad leading dimension / halo");
  const size_t n = (size_t)s.ni * s.nj;
  if (ld == s.ni && halo == 0) {
    if (s.pitch == s.ni) return hip_check(hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, st), "download");
    if (staging) {
      hipLaunchKernelGGL(k_repack, dim3((unsigned)((s.ni + 255) / 256), (unsigned)s.nj), dim3(256), 0, st, dev,
                         s.pitch, staging, s.ni, s.ni);
      int rc = hip_check(hipGetLastError(), "repack");
      if (rc) return rc;
      return hip_check(hipMemcpyAsync(host, staging, n * 8, hipMemcpyDeviceToHost, st), "download");
    }
  }
  double *p = host + (size_t)halo + (size_t)halo * ld;
  return hip_check(hipMemcpy2DAsync(p, (size_t)ld * 8, dev, (size_t)s.pitch * 8, (size_t)s.ni * 8, (size_t)s.nj,
                                    hipMemcpyDeviceToHost, st),
                   "download");
}

// P[0] = primitives of the stored state U^n (calculate_derived_quantities),
// as every stage leaves them for the next; used after cato_set_conserved.
int derive_primitives(hipStream_t st, const Solve &s) {
  FieldGeom g = s.fgeom();
  double *const *U = s.U[s.cur];
  double *const *P = s.P[0];
  double *cs = s.U[(s.cur + 1) % 3][0];
  hipLaunchKernelGGL(k_derived, dim3((unsigned)((s.ni + 255) / 256), (unsigned)s.nj), dim3(256), 0, st, g, U[0], U[1],
                     U[2], U[3], P[1], P[2], P[3], cs);
  int rc = hip_check(hipGetLastError(), "derive primitives");
  if (rc) return rc;
  rc = hip_check(hipMemcpyAsync(P[0], U[0], s.field_elems() * sizeof(double), hipMemcpyDeviceToDevice, st),
                 "derive primitives");
  if (rc) return rc;
  return fill_outflow_column(st, s, P);
}

int refresh_prims(hipStream_t st, Solve &s) {
  if (!s.prims_stale) return CATO_OK;
  double *const *U = s.U[s.cur];
  double *const *P = s.P[0];
  hipLaunchKernelGGL(k_prims_fast, dim3((unsigned)((s.ni + 255) / 256), (unsigned)s.nj), dim3(256), 0, st, s.fgeom(),
                     U[0], U[1], U[2], U[3], P[0], P[1], P[2], P[3]);
  int rc = hip_check(hipGetLastError(), "recover primitives");
  if (rc) return rc;
  s.prims_stale = false;
  return fill_outflow_column(st, s, P);
}

int fill_outflow_column(hipStream_t st, const Solve &s, double *const *P) {
  if (s.cfg.bc_plus_x != BC_OUTFLOW) return CATO_OK;
  hipLaunchKernelGGL(k_outflow_column, dim3((unsigned)((s.nj + 255) / 256)), dim3(256), 0, st, s.fgeom(), P[0], P[1],
                     P[2], P[3]);
  return hip_check(hipGetLastError(), "outflow ghost column");
}

// Every pointer a stage launch will dereference must be set (checked before launch).
int validate_stage_args(const StageArgs &s) {
  auto need = [&](const double *const *p, const char *what) -> int {
    for (int v = 0; v < 4; ++v)
      if (!p[v]) return fail(CATO_ERR_INVALID, std::string("internal: stage launch without ") + what);
    return CATO_OK;
  };
  // every set's four variables sit fstride elements apart (the kernel addresses them from p[0])
  auto strided = [&](const double *const *p, const char *what) -> int {
    if (!p[0]) return CATO_OK;
    for (int v = 1; v < 4; ++v)
      if (p[v] != p[0] + (size_t)v * (size_t)s.fstride)
        return fail(CATO_ERR_INVALID, std::string("internal: ") + what + " are not one strided set");
    return CATO_OK;
  };
  int rc;
  if ((rc = strided(s.uk, "U_k arrays")) || (rc = strided(s.u0, "U_0 arrays")) ||
      (rc = strided(s.prim, "primitive arrays")) || (rc = strided(s.out, "output arrays")) ||
      (rc = strided(s.prim_out, "output primitive arrays")) || (rc = strided(s.conv_first, "convergence arrays")))
    return rc;
  if ((rc = need(s.out, "output state"))) return rc;
  if ((rc = need(s.prim, "primitive arrays"))) return rc;
  if (s.mode != MODE_RESIDUAL && (rc = need(s.prim_out, "output primitive arrays"))) return rc;
  if (s.mode == MODE_FIRST && (rc = need(s.uk, "U_k state (first stage combine)"))) return rc;
  if ((s.mode == MODE_COMBINE || s.mode == MODE_BASE) && (rc = need(s.u0, "U_0 state"))) return rc;
  if (s.mode == MODE_COMBINE && (rc = need(s.uk, "U_k state"))) return rc;
  if (!s.dxi || !s.lxi || !s.dyj || !s.lyj) return fail(CATO_ERR_INVALID, "internal: stage launch without geometry");
  if (s.last_stage && (!s.block_min || !s.block_err))
    return fail(CATO_ERR_INVALID, "internal: last stage without reduction buffers");
  if (s.tiles_x <= 0 || s.tiles_y <= 0 || s.ni < s.nh || s.nj < s.nh || s.pitch < s.ni)
    return fail(CATO_ERR_INVALID, "internal: inconsistent stage geometry");
  return CATO_OK;
}

}  // namespace cato

using namespace cato;

struct cato_ctx {
  Solve s;
  hipStream_t stream = nullptr;
  char *mem = nullptr;
  int device = 0;
  bool timing = false;
  double stage_ms = 0.0, other_ms = 0.0;
  long long stage_n = 0, other_n = 0;
  // convergence diagnostics: one device allocation (block sums, norms, history
  // ring), the rows collected so far and the step counters of the rows
  double *diag = nullptr;
  std::vector<double> history;
  long long diag_iter = 0;
  // fluid_t%time (the host copy; after cato_advance it is current only in
  // loop->master_time until read back) and the time the solve started from
  double t_fluid = 0.0, t0 = 0.0;
  bool t_on_device = false;
  // the clock of the flux solver the first SSP stage calls (it gains the dt of
  // every attempt, retried ones included; the fluid time only the kept one's)
  double t_solver = 0.0;
  // subcycle_check: retry a cato_step with dt / 10 while a norm exceeds the tolerance
  bool subcycle = false;
  double subcycle_tol = 1e-3;
  double last_dt = 0.0;
  // cato_set_loop_dt: main.f90:140-158 dt choices of cato_run besides constant_dt
  double loop_initial_dt = 0.0, loop_startup_period = 0.0, loop_startup_dt = 0.0;
  int loop_startup = 0, loop_exact = 0;
  long long loop_iter0 = 0;
  bool hist_on = false;  // the caller asked for convergence rows (cato_set_diagnostics)
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<std::pair<hipEvent_t, hipEvent_t>, int>> pending;  // (start, stop), is_stage
  // kGraphSteps steps of the device-side loop, captured once and replayed by
  // cato_run / cato_advance while timing is off; dropped whenever the stage
  // arguments change (energy source, diagnostics)
  hipGraphExec_t steps_exec = nullptr;
  bool use_graph = true;
};

constexpr long long kGraphSteps = 32;

static void drop_steps_graph(cato_ctx *c) {
  if (c->steps_exec) (void)hipGraphExecDestroy(c->steps_exec);
  c->steps_exec = nullptr;
}

namespace {

hipEvent_t take_event(cato_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void collect_timing(cato_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.first.first, p.first.second) == hipSuccess) {
      if (p.second) {
        c->stage_ms += ms;
        c->stage_n++;
      } else {
        c->other_ms += ms;
        c->other_n++;
      }
    }
    c->ev_pool.push_back(p.first.first);
    c->ev_pool.push_back(p.first.second);
  }
  c->pending.clear();
}

template <typename F> int timed(cato_ctx *c, bool is_stage, F &&launch) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->timing) {
    a = take_event(c);
    b = take_event(c);
    if (a && b) (void)hipEventRecord(a, c->stream);
  }
  launch();
  int rc = hip_check(hipGetLastError(), "kernel launch");
  if (c->timing && a && b) {
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({{a, b}, is_stage ? 1 : 0});
    if (c->pending.size() > 4096) {
      (void)hipEventSynchronize(b);
      collect_timing(c);
    }
  }
  return rc;
}

// Arguments of a pressure-input pass: mode 0 evaluates at the fluid time, mode 1
// at the boundary time of stage m of a step (dt: the host dt when loop is null).
PressureInput pressure_args(const cato_ctx *c, int mode, int m, const LoopState *loop, double dt) {
  const Solve &s = c->s;
  PressureInput a{};
  a.tab = s.pi_n ? s.pi_tab : nullptr;
  a.n = s.pi_n;
  a.p_const = s.cfg.bc_pressure;
  a.rho_const = s.cfg.bc_density;
  a.mode = mode;
  const bool multistage = s.cfg.integrator != CATO_RK_SSP2 && s.cfg.integrator != CATO_RK_SSP3;
  a.first_multistage = (mode == 1 && multistage && m == 0) ? 1 : 0;
  a.t_at = (mode == 1 && !multistage && m == 0) ? c->t_solver : c->t_fluid;
  a.dt_at = dt;
  a.t0 = c->t0;
  a.loop = loop;
  return a;
}

}  // namespace

namespace cato {

// The pressure-input pass of stage m of a step for a solve that its device-side
// loop advances from fluid time 0 (batched and served solves): the pass of
// launch_stages with the clocks of a fresh context.
PressurePass stage_pressure_pass(const Solve &s, const StageArgs &a, int m) {
  PressurePass q{};
  q.g = s.fgeom();
  PressureInput &p = q.pa;
  p.tab = s.pi_n ? s.pi_tab : nullptr;
  p.n = s.pi_n;
  p.p_const = s.cfg.bc_pressure;
  p.rho_const = s.cfg.bc_density;
  p.mode = 1;
  const bool multistage = s.cfg.integrator != CATO_RK_SSP2 && s.cfg.integrator != CATO_RK_SSP3;
  p.first_multistage = (multistage && m == 0) ? 1 : 0;
  p.t_at = 0.0;
  p.dt_at = 0.0;
  p.t0 = 0.0;
  p.loop = a.loop;
  q.recover = (s.cfg.precision != CATO_PRECISION_FP64 && a.mode != MODE_RESIDUAL) ? 1 : 0;
  q.src = q.recover ? a.uk[0] : a.prim[0];
  q.dst = const_cast<double *>(a.prim[0]);
  q.fs = a.fstride;
  return q;
}

// A pressure-input table (initialize_1d / check_inputs, linear_interpolation_module.f90:451-478,
// 1743-1810): n = 0 (the constants) or n >= 2 strictly increasing times.
int check_pressure_table(const double *time, const double *pressure, const double *density, int n) {
  if (n != 0 && (!time || !pressure || !density)) return fail(CATO_ERR_INVALID, "null table");
  if (n == 1 || n < 0) return fail(CATO_ERR_INVALID, "cannot use linear interpolation for only one point");
  for (int k = 1; k < n; ++k)
    if (!(time[k] > time[k - 1])) return fail(CATO_ERR_INVALID, "pressure input times are not strictly increasing");
  return CATO_OK;
}

// The constant inputs are checked where the reference checks them: when a
// boundary fill uses them (pressure_input_bc.f90:204-206, 227-229).
int check_pressure_constants(const Solve &s) {
  if (s.cfg.bc_plus_x != BC_PRESSURE_INPUT || s.pi_n) return CATO_OK;
  if (!(s.cfg.bc_pressure >= 2.2250738585072014e-308))
    return fail(CATO_ERR_INVALID, "Error in pressure_input_bc_t%get_desired_pressure: desired_pressure < tiny(1.0_rk)");
  if (!(s.cfg.bc_density >= 2.2250738585072014e-308))
    return fail(CATO_ERR_INVALID, "Error in pressure_input_bc_t%get_desired_density: desired_density < tiny(1.0_rk)");
  return CATO_OK;
}

}  // namespace cato

namespace {

// Ghost column ni+1 of the primitives a stage reads, for a +x pressure input:
// from those primitives (FP64 form, residual) or recovered from the stage's
// input state (precision form).
int launch_pressure_column(cato_ctx *c, const StageArgs &a, const PressureInput &pa) {
  Solve &s = c->s;
  FieldGeom g = s.fgeom();
  double *d[4];
  for (int v = 0; v < 4; ++v) d[v] = const_cast<double *>(a.prim[v]);
  const bool recover = s.cfg.precision != CATO_PRECISION_FP64 && a.mode != MODE_RESIDUAL;
  return timed(c, false, [&] {
    if (recover)
      hipLaunchKernelGGL(k_pressure_column<true>, dim3(1), dim3(1024), 0, c->stream, g, pa, a.uk[0], a.uk[1],
                         a.uk[2], a.uk[3], d[0], d[1], d[2], d[3]);
    else
      hipLaunchKernelGGL(k_pressure_column<false>, dim3(1), dim3(1024), 0, c->stream, g, pa, a.prim[0], a.prim[1],
                         a.prim[2], a.prim[3], d[0], d[1], d[2], d[3]);
  });
}

int launch_stages(cato_ctx *c, const std::vector<StageArgs> &st) {
  Solve &s = c->s;
  dim3 grid((unsigned)(s.tiles_x * s.tiles_y));
  const bool pin = s.cfg.bc_plus_x == BC_PRESSURE_INPUT;
  int m = 0;
  for (const StageArgs &a : st) {
    int rc = validate_stage_args(a);
    if (rc) return rc;
    if (pin) {
      if ((rc = check_pressure_constants(s))) return rc;
      if ((rc = launch_pressure_column(c, a, pressure_args(c, 1, m, a.loop, a.dt_value)))) return rc;
    }
    ++m;
    rc = timed(c, true, [&] { s.launcher(grid, c->stream, a); });
    if (rc) return rc;
    if (a.mode == MODE_RESIDUAL) continue;
    if (s.cfg.precision == CATO_PRECISION_FP64) {
      if ((rc = fill_outflow_column(c->stream, s, a.prim_out))) return rc;
    } else {
      s.prims_stale = true;  // the precision form's stages store no primitives
    }
  }
  return CATO_OK;
}

int launch_reduce(cato_ctx *c, int mode, bool with_conv) {
  Solve &s = c->s;
  int nb = s.tiles_x * s.tiles_y;
  return timed(c, false, [&] {
    // the convergence sums belong to a step (modes 0 and 2 after stages), not to a CFL pass
    ConvOut cv = with_conv ? s.conv_out() : ConvOut{};
    hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(1024), 0, c->stream, nb, s.block_min, s.block_err, s.result,
                       s.loop, mode, cv);
  });
}

// Current CFL minimum and sanity flags of the stored state into result[0..1].
int launch_cfl(cato_ctx *c, int loop_mode) {
  Solve &s = c->s;
  FieldGeom g = s.fgeom();
  int nb = s.tiles_x * s.tiles_y;
  // P[0] holds the primitives of U^n (the caller's after cato_set_primitive)
  int rc = refresh_prims(c->stream, s);
  if (rc) return rc;
  rc = timed(c, false, [&] {
    hipLaunchKernelGGL(k_cfl_partial, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, g, 1, s.P[0][0], s.P[0][1],
                       s.P[0][2], s.P[0][3], s.block_min, s.block_err);
  });
  if (rc) return rc;
  return launch_reduce(c, loop_mode, false);
}

int check_ctx(cato_ctx *c) { return c ? CATO_OK : fail(CATO_ERR_INVALID, "null context"); }

// The host copy of the fluid time, read back after cato_advance left it on the device.
int sync_fluid_time(cato_ctx *c) {
  if (!c->t_on_device) return CATO_OK;
  LoopState H{};
  HIPCK(hipMemcpyAsync(&H, c->s.loop, sizeof(H), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  c->t_fluid = H.master_time;
  c->t_solver = H.master_time;
  c->t_on_device = false;
  return CATO_OK;
}

// Host <-> device copies of interior fields with (ld, halo) host layout.
int upload_field(cato_ctx *c, double *dst, const double *src, int ld, int halo, double *staging) {
  return field_up(c->stream, c->s, dst, src, ld, halo, staging);
}
int download_field(cato_ctx *c, double *dst, const double *src, int ld, int halo, double *staging) {
  if (!dst) return CATO_OK;
  return field_down(c->stream, c->s, dst, src, ld, halo, staging);
}

dim3 field_grid(const Solve &s) { return dim3((unsigned)((s.ni + 255) / 256), (unsigned)s.nj); }

template <int LIMC> void launch_interp(dim3 g, hipStream_t st, int prec, FieldGeom fg, int ld, const double *q,
                                              double *ie, double *je) {
  if (prec == CATO_PRECISION_FP32_FLUX)
    hipLaunchKernelGGL((k_interpolate<f32x, LIMC>), g, dim3(256), 0, st, fg, ld, q, ie, je);
  else
    hipLaunchKernelGGL((k_interpolate<double, LIMC>), g, dim3(256), 0, st, fg, ld, q, ie, je);
}

template <int F> void launch_flux(dim3 g, hipStream_t st, int prec, FieldGeom fg, int ld, const double *e,
                                         const double *p, int lm, double *fi, double *fj) {
  if (prec == CATO_PRECISION_FP32_FLUX)
    hipLaunchKernelGGL((k_edge_flux<f32x, F>), g, dim3(256), 0, st, fg, ld, e, p, lm, fi, fj);
  else
    hipLaunchKernelGGL((k_edge_flux<double, F>), g, dim3(256), 0, st, fg, ld, e, p, lm, fi, fj);
}


}  // namespace

extern "C" {

const char *cato_last_error(void) { return g_last_error.c_str(); }
int cato_abi_version(void) { return CATO_HIP_ABI_VERSION; }

int cato_device_bytes(const cato_config *cfg, const cato_geom *geom, unsigned long long *bytes) {
  if (!cfg || !geom || !bytes) return fail(CATO_ERR_INVALID, "null argument");
  Solve s;
  int rc = s.init(*cfg, *geom, true);
  if (rc) return rc;
  *bytes = (unsigned long long)s.device_bytes();
  return CATO_OK;
}
int cato_num_tiles(void) { return kNumTiles; }
int cato_tile_shape(int tile, int *tx, int *ty) {
  if (tile < 0 || tile >= kNumTiles) return fail(CATO_ERR_INVALID, "unknown tile id");
  if (tx) *tx = kTiles[tile].tx;
  if (ty) *ty = kTiles[tile].ty;
  return CATO_OK;
}

int cato_ctx_create(const cato_config *cfg, const cato_geom *geom, cato_ctx **out) {
  if (!cfg || !geom || !out) return fail(CATO_ERR_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CATO_ERR_NO_DEVICE, "no HIP device");
  cato_ctx *c = new cato_ctx();
  int rc = c->s.init(*cfg, *geom, true);
  if (rc) {
    delete c;
    return rc;
  }
  (void)hipGetDevice(&c->device);
  if ((rc = noalloc_scratch_init(c->device))) {
    delete c;
    return rc;
  }
  if ((rc = hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream"))) {
    delete c;
    return rc;
  }
  size_t bytes = c->s.device_bytes();
  if ((rc = hip_check(hipMalloc(&c->mem, bytes), "device allocation"))) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  c->s.bind(c->mem);
  if ((rc = c->s.upload_geometry(c->stream)) || (rc = hip_check(hipStreamSynchronize(c->stream), "sync"))) {
    cato_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return CATO_OK;
}

int cato_ctx_destroy(cato_ctx *c) {
  if (!c) return CATO_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.first.first);
    (void)hipEventDestroy(p.first.second);
  }
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  drop_steps_graph(c);
  if (c->s.src_E) (void)hipFree(c->s.src_E);
  if (c->s.pi_tab) (void)hipFree(c->s.pi_tab);
  if (c->diag) (void)hipFree(c->diag);
  if (c->mem) (void)hipFree(c->mem);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CATO_OK;
}

int cato_set_primitive(cato_ctx *c, const double *rho, const double *u, const double *v, const double *p, int ld,
                       int halo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  double *stg = s.U[(s.cur + 1) % 3][0];
  double **P = s.P[0];
  if ((rc = upload_field(c, P[0], rho, ld, halo, stg)) || (rc = upload_field(c, P[1], u, ld, halo, stg)) ||
      (rc = upload_field(c, P[2], v, ld, halo, stg)) || (rc = upload_field(c, P[3], p, ld, halo, stg)))
    return rc;
  FieldGeom g = s.fgeom();
  hipLaunchKernelGGL(k_prim_to_cons, field_grid(s), dim3(256), 0, c->stream, g, P[0], P[1], P[2], P[3],
                     s.U[s.cur][0], s.U[s.cur][1], s.U[s.cur][2], s.U[s.cur][3]);
  HIPCK(hipGetLastError());
  if ((rc = fill_outflow_column(c->stream, s, P))) return rc;
  s.prims_valid = true;
  s.prims_stale = false;
  c->diag_iter = 0;
  c->t_fluid = c->t0;
  c->t_solver = c->t0;
  c->t_on_device = false;
  return hip_check(hipStreamSynchronize(c->stream), "set_primitive");
}

int cato_set_conserved(cato_ctx *c, const double *r, const double *ru, const double *rv, const double *rE, int ld,
                       int halo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  double **U = s.U[s.cur];
  double *stg = s.U[(s.cur + 1) % 3][0];
  if ((rc = upload_field(c, U[0], r, ld, halo, stg)) || (rc = upload_field(c, U[1], ru, ld, halo, stg)) ||
      (rc = upload_field(c, U[2], rv, ld, halo, stg)) || (rc = upload_field(c, U[3], rE, ld, halo, stg)))
    return rc;
  if ((rc = derive_primitives(c->stream, s))) return rc;
  s.prims_valid = false;
  s.prims_stale = false;
  c->diag_iter = 0;
  c->t_fluid = c->t0;
  c->t_solver = c->t0;
  c->t_on_device = false;
  return hip_check(hipStreamSynchronize(c->stream), "set_conserved");
}

int cato_set_energy_source(cato_ctx *c, const double *q, int ld, int halo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  HIPCK(hipStreamSynchronize(c->stream));
  drop_steps_graph(c);  // the stage arguments carry the source pointer
  if (!q) {
    // the stages already enqueued keep reading the old field: free it after them
    if (s.src_E) {
      if ((rc = hip_check(hipStreamSynchronize(c->stream), "set_energy_source"))) return rc;
      (void)hipFree(s.src_E);
      s.src_E = nullptr;
    }
    return CATO_OK;
  }
  if (halo < 0 || ld < s.ni + 2 * halo) return fail(CATO_ERR_INVALID, "bad leading dimension / halo");
  if (!s.src_E && (rc = hip_check(hipMalloc((void **)&s.src_E, s.field_elems() * sizeof(double)), "alloc source")))
    return rc;
  if ((rc = upload_field(c, s.src_E, q, ld, halo, s.U[(s.cur + 1) % 3][0]))) return rc;
  return hip_check(hipStreamSynchronize(c->stream), "set_energy_source");
}

int cato_set_pressure_input(cato_ctx *c, const double *time, const double *pressure, const double *density, int n) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  if (s.cfg.bc_plus_x != BC_PRESSURE_INPUT) return fail(CATO_ERR_INVALID, "no pressure_input boundary at +x");
  if ((rc = check_pressure_table(time, pressure, density, n))) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  drop_steps_graph(c);  // the pressure-input passes carry the table
  if (s.pi_tab) (void)hipFree(s.pi_tab);
  s.pi_tab = nullptr;
  s.pi_n = 0;
  if (n == 0) return CATO_OK;
  std::vector<double> tab((size_t)3 * n);
  std::copy(time, time + n, tab.begin());
  std::copy(pressure, pressure + n, tab.begin() + n);
  std::copy(density, density + n, tab.begin() + 2 * n);
  if ((rc = hip_check(hipMalloc((void **)&s.pi_tab, tab.size() * sizeof(double)), "alloc pressure input"))) return rc;
  s.pi_n = n;
  HIPCK(hipMemcpyAsync(s.pi_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return hip_check(hipStreamSynchronize(c->stream), "set_pressure_input");
}

int cato_set_time(cato_ctx *c, double time) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  drop_steps_graph(c);  // the pressure-input passes carry the start time
  c->t0 = time;
  c->t_fluid = time;
  c->t_solver = time;
  c->t_on_device = false;
  return CATO_OK;
}

int cato_get_conserved(cato_ctx *c, double *r, double *ru, double *rv, double *rE, int ld, int halo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  double **U = s.U[s.cur];
  double *stg = s.U[(s.cur + 1) % 3][0];
  if ((rc = download_field(c, r, U[0], ld, halo, stg)) || (rc = download_field(c, ru, U[1], ld, halo, stg)) ||
      (rc = download_field(c, rv, U[2], ld, halo, stg)) || (rc = download_field(c, rE, U[3], ld, halo, stg)))
    return rc;
  return hip_check(hipStreamSynchronize(c->stream), "get_conserved");
}

int cato_get_primitive(cato_ctx *c, double *rho, double *u, double *v, double *p, double *cs, int ld, int halo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  FieldGeom g = s.fgeom();
  if ((rc = refresh_prims(c->stream, s))) return rc;
  // P[0]: primitives of U^n; scratch: the two non-current state buffers
  double **P = s.P[0];
  double **T = s.U[(s.cur + 1) % 3];
  hipLaunchKernelGGL(k_sound_speed, field_grid(s), dim3(256), 0, c->stream, g, P[0], P[3], T[1]);
  HIPCK(hipGetLastError());
  double *stg = T[0];
  if ((rc = download_field(c, rho, P[0], ld, halo, stg)) || (rc = download_field(c, u, P[1], ld, halo, stg)) ||
      (rc = download_field(c, v, P[2], ld, halo, stg)) || (rc = download_field(c, p, P[3], ld, halo, stg)) ||
      (rc = download_field(c, cs, T[1], ld, halo, stg)))
    return rc;
  return hip_check(hipStreamSynchronize(c->stream), "get_primitive");
}

int cato_compute_dt(cato_ctx *c, double *dt) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = launch_cfl(c, 0))) return rc;
  double r[2];
  HIPCK(hipMemcpyAsync(r, c->s.result, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  if (dt) *dt = r[0];
  return CATO_OK;
}

int cato_compute_dt_noalloc(cato_ctx *c, const double *rho, const double *u, const double *v, const double *p,
                            int ld, int halo, double *dt_device) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const Solve &s = c->s;
  if (!rho || !u || !v || !p || !dt_device) return fail(CATO_ERR_INVALID, "null device pointer");
  if (halo < 0 || ld < s.ni + 2 * halo) return fail(CATO_ERR_INVALID, "bad leading dimension / halo");
  unsigned char *base = noalloc_scratch(c->device);
  if (!base) return fail(CATO_ERR_INVALID, "allocation-free scratch is not set up on this context's device");
  const size_t n = (size_t)s.ni * s.nj;
  const size_t want = (n + kThreads - 1) / kThreads;
  // k_cfl_partial strides over the cells, so capping the workgroups at the
  // scratch's kNoAllocBlocks partials changes who reduces which cell, not the
  // minimum (min is exact in any order)
  const int nb = (int)std::min(want, (size_t)kNoAllocBlocks);
  double *bmin = reinterpret_cast<double *>(base + kNoAllocMinOff);
  int *berr = reinterpret_cast<int *>(base + kNoAllocErrOff);
  double *res = reinterpret_cast<double *>(base + kNoAllocResOff);
  FieldGeom g = s.fgeom();
  g.pitch = ld;  // the caller's layout: interior (1, 1) at halo + halo * ld
  const size_t off = (size_t)halo + (size_t)halo * (size_t)ld;
  hipLaunchKernelGGL(k_cfl_partial, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, g, 1, rho + off, u + off,
                     v + off, p + off, bmin, berr);
  HIPCK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_finalize, dim3(1), dim3(1024), 0, c->stream, nb, bmin, berr, res, (LoopState *)nullptr,
                     0, ConvOut{});
  HIPCK(hipGetLastError());
  return hip_check(hipMemcpyAsync(dt_device, res, sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                   "compute_dt_noalloc");
}

int cato_step(cato_ctx *c, double dt, int *error_code) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  if (!(dt >= 2.2250738585072014e-308))
    return fail(CATO_ERR_INVALID, "The timestep dt is < tiny(1.0_rk)");  // slau_solver.fypp:162-166
  if ((rc = sync_fluid_time(c))) return rc;
  const bool ssp = s.cfg.integrator == CATO_RK_SSP2 || s.cfg.integrator == CATO_RK_SSP3;
  // ssp_rk_2_2 / ssp_rk_3_3 with subcycle_check (mod_fluid.f90:816-823,856-887): attempts keep U^n and
  // P[0] intact, and the kept one's buffers become U^n's; rk_multistage has no retry loop
  const bool sub = c->subcycle && ssp;
  double r[2], nrm[4];
  for (;;) {
    std::vector<StageArgs> st = s.step_stages(nullptr, dt, sub);
    if ((rc = launch_stages(c, st))) return rc;
    if (ssp) c->t_solver = c->t_solver + dt;  // the first stage's solver: time + dt per solve call
    s.prims_valid = false;
    if ((rc = launch_reduce(c, 0, true))) return rc;
    HIPCK(hipMemcpyAsync(r, s.result, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    if (c->diag) HIPCK(hipMemcpyAsync(nrm, s.conv_norms, sizeof(nrm), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (!sub) break;
    bool retry = false;
    for (int v = 0; v < 4; ++v) retry = retry || nrm[v] > c->subcycle_tol;
    if (!retry) {
      const int kept = s.cfg.integrator == CATO_RK_SSP2 ? (s.cur + 2) % 3 : (s.cur + 1) % 3;
      const int kept_p = s.cfg.integrator == CATO_RK_SSP2 ? 2 : 1;
      s.cur = kept;
      for (int v = 0; v < 4; ++v) std::swap(s.P[0][v], s.P[kept_p][v]);
      drop_steps_graph(c);  // a captured device loop holds the old buffer roles
      break;
    }
    dt = dt / 10.0;
  }
  // fluid_t%integrate: time + dt (ssp_rk_3_3:844), iteration + 1 (mod_fluid.f90:422), and the dt it used
  c->last_dt = dt;
  c->t_fluid = c->t_fluid + dt;
  if (c->hist_on) {
    c->diag_iter += 1;
    const double row[kHistRow] = {(double)c->diag_iter, c->t_fluid, dt, nrm[0], nrm[1], nrm[2], nrm[3]};
    c->history.insert(c->history.end(), row, row + kHistRow);
  }
  if (c->timing) collect_timing(c);
  if (error_code) {
    int f = (int)r[1];
    *error_code = (f & 1) ? 3 : (f & 2) ? 1 : (f & 4) ? 2 : 0;
  }
  return CATO_OK;
}

// One step of the device-side loop (stages with the device dt, then the
// end-of-step reduction and loop update).
static int enqueue_step(cato_ctx *c) {
  Solve &s = c->s;
  int rc = launch_stages(c, s.step_stages(s.loop, 0.0));
  if (rc) return rc;
  s.prims_valid = false;
  return launch_reduce(c, 2, true);
}

// n steps of the device-side loop.  The arguments of every step are the same
// (the primitive buffers cycle back each step), so with timing off whole runs
// of kGraphSteps steps replay one captured graph: no host work per kernel and
// shorter gaps between them.  The first step after cato_set_primitive (which
// checks the caller's primitives) and the remainder run eagerly.
static int enqueue_steps(cato_ctx *c, long long n) {
  Solve &s = c->s;
  long long k = 0;
  int rc = CATO_OK;
  if (n > 0 && s.prims_valid) {
    if ((rc = enqueue_step(c))) return rc;
    k = 1;
  }
  if (c->use_graph && !c->timing && n - k >= kGraphSteps) {
    if (!c->steps_exec) {
      hipGraph_t g = nullptr;
      HIPCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (long long q = 0; q < kGraphSteps && !rc; ++q) rc = enqueue_step(c);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : hip_check(e, "capture device-loop steps");
      }
      hipError_t ie = hipGraphInstantiate(&c->steps_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        c->steps_exec = nullptr;
        return hip_check(ie, "instantiate device-loop steps");
      }
    }
    for (; n - k >= kGraphSteps; k += kGraphSteps) HIPCK(hipGraphLaunch(c->steps_exec, c->stream));
  }
  for (; k < n; ++k)
    if ((rc = enqueue_step(c))) return rc;
  return CATO_OK;
}

static int run_impl(cato_ctx *c, double max_time, long long max_iter, double contour_dt, double constant_dt,
                    bool poll, long long fixed_steps, long long *iters, double *time, int *err) {
  Solve &s = c->s;
  if (c->subcycle && (s.cfg.integrator == CATO_RK_SSP2 || s.cfg.integrator == CATO_RK_SSP3))
    return fail(CATO_ERR_UNSUPPORTED, "subcycling (allow_subcycle) runs in host-driven steps (cato_step) only");
  int rc = sync_fluid_time(c);
  if (rc) return rc;
  // main.f90:97-102: the driver clock starts at the fluid's time
  LoopState L{};
  L.time = c->t_fluid;
  L.master_time = c->t_fluid;
  L.next_output_time = c->t_fluid + contour_dt;
  L.contour_dt = contour_dt;
  L.max_time = max_time;
  L.constant_dt = constant_dt;
  L.iteration = 0;
  L.max_iterations = max_iter;
  if (poll) {
    L.initial_dt = c->loop_initial_dt;
    L.startup = c->loop_startup;
    L.startup_period = c->loop_startup_period;
    L.startup_dt = c->loop_startup_dt;
    L.iter0 = c->loop_iter0;
    L.exact = c->loop_exact;
  }
  HIPCK(hipMemcpyAsync(s.loop, &L, sizeof(L), hipMemcpyHostToDevice, c->stream));
  if ((rc = launch_cfl(c, 1))) return rc;
  long long launched = 0;
  const long long chunk = kGraphSteps;
  static_assert(kGraphSteps < Solve::kHistCap, "a poll must find every row of its chunk in the ring");
  LoopState H{};
  while (true) {
    long long n = poll ? chunk : fixed_steps;
    if ((rc = enqueue_steps(c, n))) return rc;
    launched += n;
    if (!poll) break;
    long long seen = H.iteration;
    HIPCK(hipMemcpyAsync(&H, s.loop, sizeof(H), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> ring;
    if (c->hist_on) {
      ring.resize((size_t)Solve::kHistCap * kHistRow);
      HIPCK(hipMemcpyAsync(ring.data(), s.conv_hist, ring.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    }
    HIPCK(hipStreamSynchronize(c->stream));
    if (c->timing) collect_timing(c);
    // rows of the steps since the last poll (at most `chunk` < kHistCap of them)
    for (long long it = seen + 1; c->hist_on && it <= H.iteration; ++it) {
      const double *row = &ring[(size_t)((it - 1) % Solve::kHistCap) * kHistRow];
      c->history.push_back((double)(c->diag_iter + it));
      c->history.push_back(row[1]);
      c->history.insert(c->history.end(), row + 2, row + kHistRow);
    }
    if (H.done) break;
  }
  if (!poll) {
    c->t_on_device = true;
    return CATO_OK;
  }
  c->diag_iter += H.iteration;
  c->t_fluid = H.master_time;
  c->t_solver = H.master_time;
  if (iters) *iters = H.iteration;
  if (time) *time = H.time;
  if (err) *err = H.error;
  return CATO_OK;
}

int cato_run(cato_ctx *c, double max_time, long long max_iterations, double contour_interval_dt, double constant_dt,
             long long *iterations, double *time, int *error_code) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (max_iterations <= 0) max_iterations = 0x7fffffffLL;
  if (!(contour_interval_dt > 0.0)) contour_interval_dt = max_time;
  return run_impl(c, max_time, max_iterations, contour_interval_dt, constant_dt, true, 0, iterations, time,
                  error_code);
}

int cato_advance(cato_ctx *c, long long n) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n <= 0) return CATO_OK;
  return run_impl(c, 1e300, n, 1e300, 0.0, false, n, nullptr, nullptr, nullptr);
}

int cato_loop_status(cato_ctx *c, long long *iterations, double *time, int *error_code) {
  int rc = check_ctx(c);
  if (rc) return rc;
  LoopState H{};
  HIPCK(hipMemcpyAsync(&H, c->s.loop, sizeof(H), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  if (c->timing) collect_timing(c);
  if (iterations) *iterations = H.iteration;
  if (time) *time = H.time;
  if (error_code) *error_code = H.error;
  return CATO_OK;
}

int cato_synchronize(cato_ctx *c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  if (c->timing) collect_timing(c);
  return CATO_OK;
}

static int set_conv_buffers(cato_ctx *c, bool need);

int cato_set_diagnostics(cato_ctx *c, int enable) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  HIPCK(hipStreamSynchronize(c->stream));
  drop_steps_graph(c);  // the last stage's arguments carry the diagnostics buffers
  c->history.clear();
  c->diag_iter = 0;
  c->hist_on = enable != 0;
  return set_conv_buffers(c, c->hist_on || c->subcycle);
}

// The convergence sums, norms and history ring (one device allocation), kept
// while the diagnostics or the subcycling need them.
static int set_conv_buffers(cato_ctx *c, bool need) {
  Solve &s = c->s;
  if (!need) {
    if (c->diag) (void)hipFree(c->diag);
    c->diag = nullptr;
    s.conv_sums = s.conv_norms = s.conv_hist = nullptr;
    return CATO_OK;
  }
  if (c->diag) return CATO_OK;
  const size_t nb = (size_t)s.tiles_x * s.tiles_y;
  const size_t n = 8 * nb + 4 + (size_t)Solve::kHistCap * kHistRow;
  HIPCK(hipMalloc((void **)&c->diag, n * sizeof(double)));
  HIPCK(hipMemsetAsync(c->diag, 0, n * sizeof(double), c->stream));
  s.conv_sums = c->diag;
  s.conv_norms = c->diag + 8 * nb;
  s.conv_hist = s.conv_norms + 4;
  return hip_check(hipStreamSynchronize(c->stream), "set_diagnostics");
}

int cato_set_subcycle(cato_ctx *c, int enable, double tolerance) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (enable && !(tolerance > 0.0)) return fail(CATO_ERR_INVALID, "subcycle_error_tolerance must be > 0");
  HIPCK(hipStreamSynchronize(c->stream));
  drop_steps_graph(c);
  c->subcycle = enable != 0;
  if (c->subcycle) c->subcycle_tol = tolerance;
  return set_conv_buffers(c, c->hist_on || c->subcycle);
}

int cato_set_loop_dt(cato_ctx *c, double initial_dt, int enable_startup, double startup_period, double startup_dt,
                     long long start_iteration, int exact_landing) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (enable_startup && !(startup_dt > 0.0)) return fail(CATO_ERR_INVALID, "startup_delta_t must be > 0");
  if (start_iteration < 0) return fail(CATO_ERR_INVALID, "start_iteration must be >= 0");
  c->loop_initial_dt = initial_dt > 0.0 ? initial_dt : 0.0;
  c->loop_startup = enable_startup != 0;
  c->loop_startup_period = startup_period;
  c->loop_startup_dt = startup_dt;
  c->loop_iter0 = start_iteration;
  c->loop_exact = exact_landing != 0;
  return CATO_OK;
}

int cato_last_dt(cato_ctx *c, double *dt) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!dt) return fail(CATO_ERR_INVALID, "null output pointer");
  *dt = c->last_dt;
  return CATO_OK;
}

int cato_get_convergence(cato_ctx *c, double *norms) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!norms) return fail(CATO_ERR_INVALID, "null output pointer");
  if (!c->diag) return fail(CATO_ERR_INVALID, "convergence diagnostics are off (cato_set_diagnostics)");
  HIPCK(hipMemcpyAsync(norms, c->s.conv_norms, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return hip_check(hipStreamSynchronize(c->stream), "get_convergence");
}

int cato_get_history(cato_ctx *c, double *rows, long long max_rows, long long *n_rows) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const long long pending = (long long)(c->history.size() / kHistRow);
  if (!rows) {
    if (n_rows) *n_rows = pending;
    return CATO_OK;
  }
  const long long n = std::min(pending, std::max(max_rows, 0LL));
  std::copy(c->history.begin(), c->history.begin() + n * kHistRow, rows);
  c->history.erase(c->history.begin(), c->history.begin() + n * kHistRow);
  if (n_rows) *n_rows = n;
  return CATO_OK;
}

int cato_set_timing(cato_ctx *c, int enable) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  collect_timing(c);
  c->timing = enable != 0;
  c->stage_ms = c->other_ms = 0.0;
  c->stage_n = c->other_n = 0;
  return CATO_OK;
}

int cato_set_device_graph(cato_ctx *c, int enable) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  drop_steps_graph(c);
  c->use_graph = enable != 0;
  return CATO_OK;
}

int cato_get_timing(cato_ctx *c, double *stage_ms, long long *stage_n, double *other_ms, long long *other_n) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (stage_ms) *stage_ms = c->stage_ms;
  if (stage_n) *stage_n = c->stage_n;
  if (other_ms) *other_ms = c->other_ms;
  if (other_n) *other_n = c->other_n;
  return CATO_OK;
}

// ---- operator-level entry points -------------------------------------------
namespace {
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

int cato_apply_bc(cato_ctx *c, double *rho, double *u, double *v, double *p, int ld) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  if (!rho || !u || !v || !p) return fail(CATO_ERR_INVALID, "null field pointer");
  int nx = s.ni + 2 * s.nh, ny = s.nj + 2 * s.nh;
  if (ld < nx) return fail(CATO_ERR_INVALID, "ld < ni + 2*nh");
  size_t n = (size_t)ld * ny;
  DevBuf d;
  HIPCK(hipMalloc(&d.p, 4 * n * 8));
  double *D = (double *)d.p;
  double *H[4] = {rho, u, v, p};
  for (int k = 0; k < 4; ++k) HIPCK(hipMemcpyAsync(D + k * n, H[k], n * 8, hipMemcpyHostToDevice, c->stream));
  FieldGeom g = s.fgeom();
  hipLaunchKernelGGL(k_apply_bc, dim3((unsigned)((nx * ny + 255) / 256)), dim3(256), 0, c->stream, g, ld, D, D + n,
                     D + 2 * n, D + 3 * n);
  HIPCK(hipGetLastError());
  if (s.cfg.bc_plus_x == BC_PRESSURE_INPUT) {
    // last, at priority 0, over the rows the other sides filled; at the fluid time
    if ((rc = check_pressure_constants(s)) || (rc = sync_fluid_time(c))) return rc;
    hipLaunchKernelGGL(k_pressure_halo, dim3(1), dim3(1024), 0, c->stream, g, pressure_args(c, 0, 0, nullptr, 0.0),
                       ld, D, D + n, D + 2 * n, D + 3 * n);
    HIPCK(hipGetLastError());
  }
  for (int k = 0; k < 4; ++k) HIPCK(hipMemcpyAsync(H[k], D + k * n, n * 8, hipMemcpyDeviceToHost, c->stream));
  return hip_check(hipStreamSynchronize(c->stream), "apply_bc");
}

int cato_interpolate_edges(cato_ctx *c, const double *q, int ld, double *i_edges, double *j_edges) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  if (!q || !i_edges || !j_edges) return fail(CATO_ERR_INVALID, "null pointer");
  int nx = s.ni + 2 * s.nh, ny = s.nj + 2 * s.nh;
  if (ld < nx) return fail(CATO_ERR_INVALID, "ld < ni + 2*nh");
  size_t n = (size_t)ld * ny, ne = (size_t)2 * (s.ni + 1) * (s.nj + 1);
  DevBuf d;
  HIPCK(hipMalloc(&d.p, (n + 2 * ne) * 8));
  double *Q = (double *)d.p, *IE = Q + n, *JE = IE + ne;
  HIPCK(hipMemcpyAsync(Q, q, n * 8, hipMemcpyHostToDevice, c->stream));
  FieldGeom g = s.fgeom();
  dim3 grid((unsigned)((ne + 255) / 256));  // 2 * (ni+1)*(nj+1) threads
  switch (s.cfg.limiter) {
    case LIM_MINMOD: launch_interp<LIM_MINMOD>(grid, c->stream, s.cfg.precision, g, ld, Q, IE, JE); break;
    case LIM_SUPERBEE: launch_interp<LIM_SUPERBEE>(grid, c->stream, s.cfg.precision, g, ld, Q, IE, JE); break;
    case LIM_VAN_LEER: launch_interp<LIM_VAN_LEER>(grid, c->stream, s.cfg.precision, g, ld, Q, IE, JE); break;
    case LIM_MLP3: launch_interp<LIM_MLP3>(grid, c->stream, s.cfg.precision, g, ld, Q, IE, JE); break;
    default: launch_interp<LIM_MLP5>(grid, c->stream, s.cfg.precision, g, ld, Q, IE, JE); break;
  }
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(i_edges, IE, ne * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipMemcpyAsync(j_edges, JE, ne * 8, hipMemcpyDeviceToHost, c->stream));
  return hip_check(hipStreamSynchronize(c->stream), "interpolate_edges");
}

int cato_edge_fluxes(cato_ctx *c, const double *edges, const double *p, int ld, double *iflux, double *jflux) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  if (!edges || !p || !iflux || !jflux) return fail(CATO_ERR_INVALID, "null pointer");
  int nx = s.ni + 2 * s.nh, ny = s.nj + 2 * s.nh;
  if (ld < nx) return fail(CATO_ERR_INVALID, "ld < ni + 2*nh");
  size_t n = (size_t)ld * ny, ne = (size_t)2 * (s.ni + 1) * (s.nj + 1);
  size_t nfi = (size_t)4 * (s.ni + 1) * s.nj, nfj = (size_t)4 * s.ni * (s.nj + 1);
  const size_t nedge = (s.cfg.flux == FLUX_M_AUSMPW_PLUS ? 16 : 8) * ne;  // + superbee states
  DevBuf d;
  HIPCK(hipMalloc(&d.p, (nedge + n + nfi + nfj) * 8));
  double *E = (double *)d.p, *P = E + nedge, *FI = P + n, *FJ = FI + nfi;
  HIPCK(hipMemcpyAsync(E, edges, nedge * 8, hipMemcpyHostToDevice, c->stream));
  HIPCK(hipMemcpyAsync(P, p, n * 8, hipMemcpyHostToDevice, c->stream));
  FieldGeom g = s.fgeom();
  int nt = (s.ni + 1) * s.nj + s.ni * (s.nj + 1);
  dim3 grid((unsigned)((nt + 255) / 256));
  int lm = s.cfg.apply_low_mach_fix, pr = s.cfg.precision;
  switch (s.cfg.flux) {
    case FLUX_AUSMPW_PLUS: launch_flux<FLUX_AUSMPW_PLUS>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_SLAU: launch_flux<FLUX_SLAU>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_SLAU2: launch_flux<FLUX_SLAU2>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_SD_SLAU: launch_flux<FLUX_SD_SLAU>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_SD_SLAU2: launch_flux<FLUX_SD_SLAU2>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_ROE: launch_flux<FLUX_ROE>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_AUSM_PLUS: launch_flux<FLUX_AUSM_PLUS>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    case FLUX_M_AUSMPW_PLUS: launch_flux<FLUX_M_AUSMPW_PLUS>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
    default: launch_flux<FLUX_HLLC>(grid, c->stream, pr, g, ld, E, P, lm, FI, FJ); break;
  }
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(iflux, FI, nfi * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipMemcpyAsync(jflux, FJ, nfj * 8, hipMemcpyDeviceToHost, c->stream));
  return hip_check(hipStreamSynchronize(c->stream), "edge_fluxes");
}

int cato_residual(cato_ctx *c, double *rho, double *u, double *v, double *p, double *d_rho, double *d_rho_u,
                  double *d_rho_v, double *d_rho_E, int ld) {
  int rc = check_ctx(c);
  if (rc) return rc;
  Solve &s = c->s;
  double *D[4] = {d_rho, d_rho_u, d_rho_v, d_rho_E};
  for (int k = 0; k < 4; ++k)
    if (!D[k]) return fail(CATO_ERR_INVALID, "null pointer");
  // 1. boundary fill of the caller's halo'd primitives (intent(inout) in the reference)
  if ((rc = cato_apply_bc(c, rho, u, v, p, ld))) return rc;
  // 2. residual from the interior primitives (ghosts resolved in-kernel by the same BC map);
  //    scratch: the two state buffers that do not hold U^n, so a pending initial state survives
  double **Pbuf = s.U[(s.cur + 1) % 3];
  double **Obuf = s.U[(s.cur + 2) % 3];
  const double *P[4] = {rho, u, v, p};
  for (int k = 0; k < 4; ++k)
    if ((rc = upload_field(c, Pbuf[k], P[k], ld, s.nh, nullptr))) return rc;
  if ((rc = fill_outflow_column(c->stream, s, Pbuf))) return rc;
  StageArgs a = s.stage_args(MODE_RESIDUAL, 0, 0, 0, -1, -1, (s.cur + 2) % 3, 0, -1, false, nullptr, 1.0, false);
  for (int k = 0; k < 4; ++k) a.prim[k] = Pbuf[k];
  a.block_min = nullptr;
  a.block_err = nullptr;
  if ((rc = validate_stage_args(a))) return rc;
  if (s.cfg.bc_plus_x == BC_PRESSURE_INPUT &&
      (rc = launch_pressure_column(c, a, pressure_args(c, 0, 0, nullptr, 0.0))))
    return rc;
  dim3 grid((unsigned)(s.tiles_x * s.tiles_y));
  s.launcher(grid, c->stream, a);
  HIPCK(hipGetLastError());
  int ny = s.nj + 2 * s.nh;
  for (int k = 0; k < 4; ++k) {
    std::memset(D[k], 0, (size_t)ld * ny * sizeof(double));
    if ((rc = download_field(c, D[k], Obuf[k], ld, s.nh, nullptr))) return rc;
  }
  return hip_check(hipStreamSynchronize(c->stream), "residual");
}

}  // extern "C"
