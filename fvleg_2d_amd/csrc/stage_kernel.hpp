// Fused SSP-RK stage kernel for gfx950.
//
// One workgroup owns a TX x TY tile of interior cells.  In one pass it:
//   1. loads the primitive variables (rho, u, v, p) of U_k for the tile plus
//      an H-cell halo into LDS, resolving ghost cells through the boundary
//      conditions' per-axis index map (apply_primitive_bc);
//   2. reconstructs the L/R states and evaluates every edge flux of the tile
//      exactly once (muscl_*%interpolate_edge_values + get_<dir>flux);
//   3. forms dU/dt per cell (flux_split_edges) and the SSP stage combination
//      (fluid_t operators, mod_fluid.f90:593-669) and writes U_{k+1};
//   4. recovers the primitives of U_{k+1} once per cell and stores them for
//      the next stage (fluid_t%calculate_derived_quantities), and on the last
//      stage reduces the CFL time step and the sanity flags per workgroup
//      (get_timestep, sanity_check).
// FP64 form: each cell's primitives are computed once per stage instead of
// once per tile that needs them as halo; the price is 32 B/cell of extra
// writes on a kernel whose time is spent in FP64 division, not in HBM
// traffic.  Precision form (R = f32x): the balance is the other way round, so
// step 1 reads U_k and recovers the primitives of tile and halo itself, and
// step 4 stores none.  The geometry of rectilinear grids is four 1D arrays;
// general quadrilateral grids (GEOM) read per-cell lengths, volumes and edge
// normals.
#pragma once

#include "physics.hpp"


namespace cato {

constexpr int kThreads = 256;
constexpr int kNumXcd = 8;  // MI355X: 8 XCDs with private L2, workgroups dealt round-robin

// XCD-aware tile order: workgroup b runs on XCD b % 8; give each XCD a
// contiguous run of tiles so neighbouring tiles (which share halo rows) hit
// the same L2.
__device__ inline int xcd_tile(int b, int nb) {
  int x = b % kNumXcd, local = b / kNumXcd;
  int q = nb / kNumXcd, r = nb % kNumXcd;
  int start = x * q + (x < r ? x : r);
  return start + local;
}

// Ghost-cell resolution: the per-axis composition of the BC fills
// (flux_solver.f90:138-185 with periodic/zero-gradient/symmetry/outflow on one
// image).  Every fill but the +x outflow copies a (possibly mirrored) cell, so
// the fills compose per axis; the +x outflow has the lowest priority and is
// applied first among equals, so its transform acts last on the mapped cell.
__device__ inline int map_axis(int g, int n, int bc_hi, int bc_lo, bool &flip) {
  flip = false;
  if (g < 1) {
    if (bc_lo == BC_PERIODIC) return g + n;
    if (bc_lo == BC_ZERO_GRADIENT || bc_lo == BC_OUTFLOW || bc_lo == BC_PRESSURE_INPUT) return 1;
    flip = true;
    return 1 - g;
  }
  if (g > n) {
    if (bc_hi == BC_PERIODIC) return g - n;
    if (bc_hi == BC_ZERO_GRADIENT || bc_hi == BC_OUTFLOW || bc_hi == BC_PRESSURE_INPUT) return n;
    flip = true;
    return 2 * n + 1 - g;
  }
  return g;
}

// outlet_bc_t%apply at +x (outlet_bc.f90:96-123): every ghost layer takes the
// last interior cell of its row, extrapolated along the outgoing
// characteristic to the ambient pressure p_b when the flow there is subsonic.
// The reference leaves the ghost untouched when rho or p_b is not > tiny; here
// it is then a plain copy (such a state fails the step's density check anyway).
__device__ inline Prim<double> outflow_ghost(Prim<double> d, double gamma, double p_b) {
  const double TINY = 2.2250738585072014e-308;
  if (!(p_b > TINY && d.rho > TINY)) return d;
  double cs = sqrt(gamma * absd(d.p / d.rho));
  double vel = sqrt(d.u * d.u + d.v * d.v);
  double mach = vel / cs;
  double nx = d.u / (vel + 1e-30), ny = d.v / (vel + 1e-30);
  if (mach > 1.0) return d;
  Prim<double> g;
  g.rho = d.rho + (p_b - d.p) / (cs * cs);
  g.u = d.u + (nx * (d.p - p_b)) / (d.rho * cs);
  g.v = d.v + (ny * (d.p - p_b)) / (d.rho * cs);
  g.p = p_b;
  return g;
}

template <typename T> __device__ inline T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = (w < v) ? w : v;
  }
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}
__device__ inline int wave_or(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// The four variables of a strided set at one cell.
__device__ inline Prim<double> load4(const double *p, size_t off, long long fstride) {
  const size_t fs = (size_t)fstride;
  return Prim<double>{p[off], p[off + fs], p[off + 2 * fs], p[off + 3 * fs]};
}

// Sub-domain of a split solve: halo'd cell (gi, gj) when a side takes its
// ghosts from the halo store.  The physical sides map first (as the reference
// fills boundary ghosts after the exchange, over the full transverse extent),
// then the cell is interior, in an exchanged strip or in an exchanged corner.
__device__ inline Prim<double> load_prim_split(const StageArgs &a, int gi, int gj) {
  const int ni = a.ni, nj = a.nj, nh = a.nh;
  const bool xh = (gi > ni && (a.halo_mask & 1)) || (gi < 1 && (a.halo_mask & 2));
  const bool yh = (gj > nj && (a.halo_mask & 4)) || (gj < 1 && (a.halo_mask & 8));
  bool fu = false, fv = false;
  const int si = xh ? gi : map_axis(gi, ni, a.bc[0], a.bc[1], fu);
  const int sj = yh ? gj : map_axis(gj, nj, a.bc[2], a.bc[3], fv);
  const bool xo = si < 1 || si > ni, yo = sj < 1 || sj > nj;
  Prim<double> pr;
  if (!xo && !yo) {
    const size_t off = (size_t)(sj - 1) * a.pitch + (size_t)(si - 1);
    pr = load4(a.prim[0], off, a.fstride);
  } else {
    const int kx = si > ni ? si - ni - 1 : si + nh - 1, ky = sj > nj ? sj - nj - 1 : sj + nh - 1;
    int region;
    size_t stride, pos;  // element (v, layer, pos) at v * stride + layer * len + pos
    size_t len;
    if (xo && !yo) {
      region = si > ni ? 0 : 1;
      len = (size_t)nj;
      pos = (size_t)kx * len + (size_t)(sj - 1);
    } else if (!xo) {
      region = sj > nj ? 2 : 3;
      len = (size_t)ni;
      pos = (size_t)ky * len + (size_t)(si - 1);
    } else {
      region = 4 + (si > ni ? 0 : 1) + (sj > nj ? 0 : 2);
      len = (size_t)nh;
      pos = (size_t)ky * len + (size_t)kx;
    }
    stride = (size_t)nh * len;
    const double *h = a.halo + halo_region_offset(region, ni, nj, nh) + pos;
    pr = Prim<double>{h[0], h[stride], h[2 * stride], h[3 * stride]};
  }
  if (fu) pr.u = -pr.u;
  if (fv) pr.v = -pr.v;
  return pr;
}

// Primitives of halo'd cell (gi, gj), with the BC mapping.  The stored
// primitives of a state are those fluid_t%calculate_derived_quantities gives
// for it (written by the stage that produced the state), or the caller's on
// the first stage after cato_set_primitive.
__device__ inline Prim<double> load_prim(const StageArgs &a, int gi, int gj) {
  if (a.halo_mask) return load_prim_split(a, gi, gj);
  bool fu, fv;
  int si = map_axis(gi, a.ni, a.bc[0], a.bc[1], fu);
  // +x outflow / pressure input: every ghost layer reads column ni+1 of the
  // primitive arrays, which the producer of the state filled from column ni.
  // outflow_ghost commutes exactly with the v negation of a symmetry fill in
  // y, so the y-mapped corners equal the reference's fill-after-mirror; the
  // pressure input ignores the domain v and writes v = +0 on every row.
  const bool gcol = gi > a.ni && ghost_column_bc(a.bc[0]);
  if (gcol) si = a.ni + 1;
  int sj = map_axis(gj, a.nj, a.bc[2], a.bc[3], fv);
  if (gcol && a.bc[0] == BC_PRESSURE_INPUT) fv = false;
  size_t off = (size_t)(sj - 1) * a.pitch + (size_t)(si - 1);
  Prim<double> pr = load4(a.prim[0], off, a.fstride);
  if (fu) pr.u = -pr.u;
  if (fv) pr.v = -pr.v;
  return pr;
}

// Cell-local state traffic of phase 3 (U_0 at the cell, the new state and its
// primitives): no other tile of the launch touches it, so the FP64 form
// streams it past L2 with nontemporal hints, leaving L2 to the halo rows that
// neighbouring tiles re-read (1.5% faster there; the precision form, which
// re-reads its input state at the cell, is better off without the hints).
template <bool NT> __device__ inline double ld_cell(const double *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT> __device__ inline void st_cell(double *p, double x) {
  if constexpr (NT) {
    __builtin_nontemporal_store(x, p);
  } else {
    *p = x;
  }
}

// Precision form: primitives of halo'd cell (gi, gj) recovered from the
// stage's input conserved state (the same reciprocal-based recovery the
// form's previous stage would have stored), BC mapping and the +x outflow
// fill included.  The form stores no primitive arrays: 64 B/cell less HBM
// traffic per stage for the price of the recovery on halo cells.
__device__ inline Prim<double> load_recovered(const StageArgs &a, int gi, int gj) {
  bool fu, fv;
  // the pressure input's ghosts need the whole column (its largest pressure):
  // a pass before the stage stores them in column ni+1 of the primitive arrays
  if (gi > a.ni && a.bc[0] == BC_PRESSURE_INPUT) {
    size_t off = (size_t)(map_axis(gj, a.nj, a.bc[2], a.bc[3], fv) - 1) * a.pitch + (size_t)a.ni;
    return load4(a.prim[0], off, a.fstride);
  }
  int si = map_axis(gi, a.ni, a.bc[0], a.bc[1], fu);
  int sj = map_axis(gj, a.nj, a.bc[2], a.bc[3], fv);
  size_t off = (size_t)(sj - 1) * a.pitch + (size_t)(si - 1);
  double cs;
  const size_t fs = (size_t)a.fstride;
  Prim<double> pr = cons_to_prim_fast(a.uk[0][off], a.uk[0][off + fs], a.uk[0][off + 2 * fs],
                                      a.uk[0][off + 3 * fs], a.gamma, a.gamma - 1.0, 0, &cs);
  if (fu) pr.u = -pr.u;
  if (fv) pr.v = -pr.v;
  if (gi > a.ni && a.bc[0] == BC_OUTFLOW) pr = outflow_ghost(pr, a.gamma, a.bc_pressure);
  return pr;
}

template <int LIMC> constexpr int halo_of() { return LIMC == LIM_MLP5 ? 3 : 2; }

// Reconstruct one variable at an edge from LDS row/column access.
// s(k) returns q at normal offset k (k=-2..3) on the edge's row; t(k, dt) the
// transverse neighbours.
template <typename R, int LIMC, class D = IeeeOps>
__device__ inline void recon_var(const R *base, int sn, int st, R &L, R &Rv) {
  // base points at cell 0 (left of the edge); sn = stride along the normal, st = transverse stride
  if constexpr (LIMC == LIM_MLP3 || LIMC == LIM_MLP5) {
    R qm3 = (LIMC == LIM_MLP5) ? base[-2 * sn] : R(0);
    R q3 = (LIMC == LIM_MLP5) ? base[3 * sn] : R(0);
    recon_mlp<R, (LIMC == LIM_MLP5 ? 5 : 3), D>(qm3, base[-sn], base[0], base[sn], base[2 * sn], q3, base[st], base[-st],
                                               base[sn + st], base[sn - st], L, Rv);
  } else {
    recon_tvd2<R, LIMC, D>(base[-sn], base[0], base[sn], base[2 * sn], L, Rv);
  }
}

// Threads of the single-solve stage kernel for a tile: 256, or one per cell on
// the 512-cell tiles (32x16, 64x8) in FP64 and in the precision form with MLP5,
// where that was measured faster (FP64 KH 2048^2 32x16: 348 -> 255 us; Riemann
// 1024^2 HLLC + MLP5 32x16: 136 -> 112 us, precision form 55.8 -> 53.1 us; the
// precision form with TVD was 5 % slower).  The batched kernel runs every tile
// with kThreads.
template <typename R, int LIMC, int TX, int TY> constexpr int stage_threads() {
  return (TX * TY == 2 * kThreads && (std::is_same<R, double>::value || LIMC == LIM_MLP5)) ? 2 * kThreads
                                                                                          : kThreads;
}

// In the single-solve kernel, FP64 tiles of one cell per thread split the edge
// pass into rounds of one edge direction (split_edge_rounds, the SPLIT layout),
// whose flux takes the direction folded at compile time.  The batched kernel,
// whose body holds every scheme, keeps one loop: three code copies per scheme
// made its 256-solve workload 6 % slower.  The precision form keeps one loop
// too: it lets the compiler fuse products into FMAs, which can round one
// formula differently in two code copies, and a tile-boundary edge, which both
// neighbouring tiles compute, must round alike in both.  With an MLP5
// reconstruction, and on the 512-cell tiles, the FP64 edge pass also keeps each
// cell's bottom j-edge flux in the registers of the thread that updates the
// cell (hold_j_flux): the j-flux array then lives in the primitives' LDS once
// the edge pass has read them, which takes an MLP5 16x16 tile's LDS from 33 KB
// to 25 KB (five workgroups per CU instead of four) and a TVD 32x16 tile's from
// 57 KB to 41 KB (three instead of two); the 256-cell TVD tiles keep the j-flux
// array (their extra barrier costs more than the occupancy gains).
template <typename R, int TX, int TY, int NT> constexpr bool split_edge_rounds() {
  return std::is_same<R, double>::value && TX * TY == NT;
}
template <int LIMC, int NT, bool SPLIT> constexpr bool hold_j_flux() {
  return SPLIT && (LIMC == LIM_MLP5 || NT > kThreads);
}

// LDS image of one tile: primitives with halo, i-edge and j-edge fluxes.  With
// hold_j_flux the j-edge fluxes below the tile's cells (sgj) overlay the
// primitives after the edge pass, and sfj holds only the row above the tile.
template <int LIMC, typename R, int TX, int TY, int NT = kThreads, bool SPLIT = false> struct TileSmem {
  static constexpr int H = halo_of<LIMC>();
  static constexpr int RX = TX + 2 * H, RY = TY + 2 * H;
  static constexpr bool kSplit = SPLIT;
  static constexpr bool kHoldJ = hold_j_flux<LIMC, NT, SPLIT>();
  static constexpr int kTopRow = kHoldJ ? 0 : TY;  // row of sfj holding the j-edges above the tile
  union {
    R sp[4][RY][RX];
    R sgj[4][TY][TX];
  };
  R sfi[4][TY][TX + 1];
  R sfj[4][kTopRow + 1][TX];
};

// Block-wide scratch for the CFL / sanity / convergence reductions and the
// per-wave votes on the FastOps operand ranges.
template <int NW> struct RedSmemT {
  double m[NW];
  int e[NW];
  int fast[NW];
  double s[NW][8];
};
using RedSmem = RedSmemT<kThreads / 64>;

// Edge (ex, ey) of a tile, an i-edge (idir) or a j-edge: its L/R
// reconstruction from the LDS primitives and its flux.  i-edge (ex, ey) lies
// between tile cells (ex - 1, ey) and (ex, ey), j-edge (ex, ey) between
// (ex, ey - 1) and (ex, ey).  Returns false for an edge outside the domain.
// KIND 0 / 1: every edge of the caller is an i- / j-edge (the flux's direction
// folded at compile time); -1: either, read from idir (direction code 4).
// D / DF: the division and square root of the FP64 reconstruction / flux
// (IeeeOps, or FastOps on in-range tiles).  GEOM: general quadrilateral cells,
// the normal read from the per-cell geometry (the right edge of its left cell
// for an i-edge, the top edge for a j-edge).
template <int FLUX, int LIMC, typename R, int TX, int TY, class D, class DF, bool GEOM, int KIND, int NT, bool SPLIT>
__device__ __forceinline__ bool tile_edge(const StageArgs &a, TileSmem<LIMC, R, TX, TY, NT, SPLIT> &sm, int i0, int j0,
                                          bool idir, int ex, int ey, R (&f)[4]) {
  constexpr int H = halo_of<LIMC>();
  constexpr int RX = TX + 2 * H;
  if constexpr (KIND >= 0) idir = KIND == 0;
  auto &sp = sm.sp;
  // edge between cell (ci, cj) and its +normal neighbour
  int ci = idir ? i0 - 1 + ex : i0 + ex;
  int cj = idir ? j0 + ey : j0 - 1 + ey;
  if (ci > a.ni || cj > a.nj) return false;
  int lx = ci - (i0 - H), ly = cj - (j0 - H);
  const int sn = idir ? 1 : RX, st = idir ? RX : 1;
  EdgeState<R> s;
  recon_var<R, LIMC, D>(&sp[0][ly][lx], sn, st, s.rL, s.rR);
  recon_var<R, LIMC, D>(&sp[1][ly][lx], sn, st, s.uL, s.uR);
  recon_var<R, LIMC, D>(&sp[2][ly][lx], sn, st, s.vL, s.vR);
  recon_var<R, LIMC, D>(&sp[3][ly][lx], sn, st, s.pL, s.pR);
  EdgeState<R> sb;  // superbee reconstruction of the same edge (M-AUSMPW+ only)
  if constexpr (needs_superbee_states<FLUX>()) {
    recon_var<R, LIM_SUPERBEE, D>(&sp[0][ly][lx], sn, st, sb.rL, sb.rR);
    recon_var<R, LIM_SUPERBEE, D>(&sp[1][ly][lx], sn, st, sb.uL, sb.uR);
    recon_var<R, LIM_SUPERBEE, D>(&sp[2][ly][lx], sn, st, sb.vL, sb.vR);
    recon_var<R, LIM_SUPERBEE, D>(&sp[3][ly][lx], sn, st, sb.pL, sb.pR);
  }
  PressNbr<R> pn;
  if constexpr (needs_pressure_nbrs<FLUX>()) {
    const R *pc = &sp[3][ly][lx];
    pn.p00 = pc[0];
    pn.p10 = pc[sn];
    pn.p20 = pc[2 * sn];
    pn.pm10 = pc[-sn];
    pn.p0p = pc[st];
    pn.p0m = pc[-st];
    pn.p1p = pc[sn + st];
    pn.p1m = pc[sn - st];
  }
  // direction code: 0 / 1 (2 / 3 with the normals of quadrilateral cells) when
  // the caller's edges are all of one kind, else 4: one flux code path for i-
  // and j-edges, so that a wave holding both kinds runs it once
  constexpr int DIR = KIND < 0 ? 4 : (GEOM ? 2 : 0) + KIND;
  s.ax = idir ? 0 : 1;
  if constexpr (GEOM) {
    const size_t gc = (size_t)(ci + a.nh - 1) + (size_t)(cj + a.nh - 1) * (size_t)a.quad_ncx;
    const size_t n = (size_t)a.quad_n;
    s.nx = (R)a.quad[(idir ? Q_NRX : Q_NTX) * n + gc];
    s.ny = (R)a.quad[(idir ? Q_NRY : Q_NTY) * n + gc];
  } else {
    s.nx = idir ? R(1) : R(0);
    s.ny = idir ? R(0) : R(1);
  }
  edge_flux<R, FLUX, DIR, DF>(gas_of<R>(a.gamma, a.gas), a.low_mach_fix, s, pn, f, &sb);
  return true;
}

// Phase 2 of a stage tile: every i- and j-edge of the tile once, its flux
// written to LDS.  SPLIT layout (one cell per thread): the left i-edge of each
// thread's cell, then the i-edges right of the tile and the j-edges above it
// (the first TY + TX threads), then the j-edge below each thread's cell, which
// with hold_j_flux stays in `hold` (zero outside the domain) and otherwise goes
// to sfj; the first and last rounds take a single edge direction, folded at
// compile time.  Otherwise the edges in order, i-edges first, with the
// direction at run time.
template <int FLUX, int LIMC, typename R, int TX, int TY, class D, class DF, bool GEOM, int NT, bool SPLIT>
__device__ inline void edge_pass(const StageArgs &a, TileSmem<LIMC, R, TX, TY, NT, SPLIT> &sm, int i0, int j0,
                                 R (&hold)[4]) {
  constexpr int NIE = (TX + 1) * TY, NJE = TX * (TY + 1);
  const int tid = threadIdx.x;
  using Smem = TileSmem<LIMC, R, TX, TY, NT, SPLIT>;
  if constexpr (SPLIT) {
    const int cy = tid / TX, cx = tid - cy * TX;
    R f[4];
    if (tile_edge<FLUX, LIMC, R, TX, TY, D, DF, GEOM, 0, NT, SPLIT>(a, sm, i0, j0, true, cx, cy, f)) {
#pragma unroll
      for (int v = 0; v < 4; ++v) sm.sfi[v][cy][cx] = f[v];
    }
    if (tid < TY + TX) {
      const bool idir = tid < TY;
      const int ex = idir ? TX : tid - TY, ey = idir ? tid : TY;
      if (tile_edge<FLUX, LIMC, R, TX, TY, D, DF, GEOM, -1, NT, SPLIT>(a, sm, i0, j0, idir, ex, ey, f)) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (idir)
            sm.sfi[v][ey][ex] = f[v];
          else
            sm.sfj[v][Smem::kTopRow][ex] = f[v];
        }
      }
    }
    if constexpr (Smem::kHoldJ) {
      if (!tile_edge<FLUX, LIMC, R, TX, TY, D, DF, GEOM, 1, NT, SPLIT>(a, sm, i0, j0, false, cx, cy, hold)) {
#pragma unroll
        for (int v = 0; v < 4; ++v) hold[v] = R(0);
      }
    } else {
      if (tile_edge<FLUX, LIMC, R, TX, TY, D, DF, GEOM, 1, NT, SPLIT>(a, sm, i0, j0, false, cx, cy, f)) {
#pragma unroll
        for (int v = 0; v < 4; ++v) sm.sfj[v][cy][cx] = f[v];
      }
    }
  } else {
    for (int it = tid; it < NIE + NJE; it += NT) {
      const bool idir = it < NIE;
      int ex, ey;
      if (idir) {
        ey = it / (TX + 1);
        ex = it - ey * (TX + 1);
      } else {
        int k = it - NIE;
        ey = k / TX;
        ex = k - ey * TX;
      }
      R f[4];
      if (!tile_edge<FLUX, LIMC, R, TX, TY, D, DF, GEOM, -1, NT, SPLIT>(a, sm, i0, j0, idir, ex, ey, f)) continue;
      R *dst = idir ? &sm.sfi[0][ey][ex] : &sm.sfj[0][ey][ex];
      const int fstride = idir ? TY * (TX + 1) : (TY + 1) * TX;
#pragma unroll
      for (int v = 0; v < 4; ++v) dst[v * fstride] = f[v];
    }
  }
}

// One tile of one stage.  Writes the block's CFL partial and sanity flags to
// a.block_min[tile] / a.block_err[tile] when requested.  OUTFLOW_COL: the tile
// also writes the +x outflow ghost column of the new state (the batched
// kernel); the single-solve launch leaves that to k_outflow_column, since the
// extra epilogue slows every tile of the fused kernel by ~5%.
template <int FLUX, int LIMC, typename R, int TX, int TY, bool OUTFLOW_COL = false, bool GEOM = false,
          int NT = kThreads, bool SPLIT = false>
__device__ inline void stage_tile(const StageArgs &a, int tile, double dt, TileSmem<LIMC, R, TX, TY, NT, SPLIT> &sm,
                                  RedSmemT<NT / 64> &red) {
  constexpr int H = halo_of<LIMC>();
  constexpr int RX = TX + 2 * H, RY = TY + 2 * H;
  // FP64 form: the reference's arithmetic bit for bit; precision form (f32x):
  // a lighter FP64 epilogue (reciprocal-based primitive recovery and flux balance)
  constexpr bool kExact = !std::is_same<R, f32x>::value;
  auto &sp = sm.sp;
  auto &sfi = sm.sfi;
  auto &sfj = sm.sfj;
  auto &red_min = red.m;
  auto &red_err = red.e;
  const int tid = threadIdx.x;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const int i0 = tx * TX + 1, j0 = ty * TY + 1;
  const int ni = a.ni, nj = a.nj, nh = a.nh;
  int err = 0;

  // ---- 1. primitives of the tile + halo ---------------------------------
  // (precision form outside MODE_RESIDUAL: recovered from U_k, and checked
  // for NaN here, since no earlier stage stored and checked them)
  const bool recover = !kExact && a.mode != MODE_RESIDUAL;
  // FP64: whether every cell the edges read keeps the FastOps operations of the
  // reconstruction and of the flux (where it has a FastOps path) in range
  // (fast_ops_cell).  TVD2 under a flux without one divides too little for the
  // range vote to pay.
  constexpr bool kFastScheme =
      kExact && (fast_ops_flux<FLUX>() || LIMC == LIM_MLP3 || LIMC == LIM_MLP5);
  bool in_range = true;
  // A tile whose halo lies inside the domain (all but the tiles along its
  // sides) reads its cells directly, without the boundary map
  const bool inner = !a.halo_mask && i0 - H >= 1 && j0 - H >= 1 && i0 + TX - 1 + H <= ni && j0 + TY - 1 + H <= nj;
  if (inner) {
    const bool check = recover || a.check_input;
    const size_t base = (size_t)(j0 - H - 1) * a.pitch + (size_t)(i0 - H - 1);
    const size_t fs = (size_t)a.fstride;
    for (int idx = tid; idx < RX * RY; idx += NT) {
      int ly = idx / RX, lx = idx - ly * RX;
      const size_t off = base + (size_t)ly * a.pitch + (size_t)lx;
      Prim<double> pr;
      if (recover) {
        double cs;
        pr = cons_to_prim_fast(a.uk[0][off], a.uk[0][off + fs], a.uk[0][off + 2 * fs], a.uk[0][off + 3 * fs],
                               a.gamma, a.gamma - 1.0, 0, &cs);
      } else {
        pr = load4(a.prim[0], off, a.fstride);
      }
      if (check && (pr.rho != pr.rho || pr.u != pr.u || pr.v != pr.v || pr.p != pr.p)) err |= 1;
      if constexpr (kFastScheme) in_range = in_range && fast_ops_cell<fast_ops_bound<FLUX, LIMC>()>(pr);
      sp[0][ly][lx] = (R)pr.rho;
      sp[1][ly][lx] = (R)pr.u;
      sp[2][ly][lx] = (R)pr.v;
      sp[3][ly][lx] = (R)pr.p;
    }
  } else {
    for (int idx = tid; idx < RX * RY; idx += NT) {
      int ly = idx / RX, lx = idx - ly * RX;
      int gi = i0 - H + lx, gj = j0 - H + ly;
      Prim<double> pr{0.0, 0.0, 0.0, 0.0};
      if (gi >= 1 - nh && gi <= ni + nh && gj >= 1 - nh && gj <= nj + nh) {
        pr = recover ? load_recovered(a, gi, gj) : load_prim(a, gi, gj);
        if ((recover || a.check_input) && gi >= 1 && gi <= ni && gj >= 1 && gj <= nj &&
            (pr.rho != pr.rho || pr.u != pr.u || pr.v != pr.v || pr.p != pr.p))
          err |= 1;
        if constexpr (kFastScheme) in_range = in_range && fast_ops_cell<fast_ops_bound<FLUX, LIMC>()>(pr);
      }
      sp[0][ly][lx] = (R)pr.rho;
      sp[1][ly][lx] = (R)pr.u;
      sp[2][ly][lx] = (R)pr.v;
      sp[3][ly][lx] = (R)pr.p;
    }
  }
  if constexpr (kFastScheme) {
    const int vote = __all(in_range ? 1 : 0);
    if ((tid & 63) == 0) red.fast[tid >> 6] = vote;
  }
  __syncthreads();

  // ---- 2. edge states and fluxes, each edge once --------------------------
  constexpr bool kHoldJ = TileSmem<LIMC, R, TX, TY, NT, SPLIT>::kHoldJ;
  R hold[4];  // hold_j_flux: the flux through the bottom edge of this thread's cell
  if constexpr (kFastScheme) {
    bool fast = a.fast_ops != 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) fast = fast && red.fast[w] != 0;
    using DF = typename std::conditional<fast_ops_flux<FLUX>(), FastOps, IeeeOps>::type;
    if (fast)
      edge_pass<FLUX, LIMC, R, TX, TY, FastOps, DF, GEOM, NT, SPLIT>(a, sm, i0, j0, hold);
    else
      edge_pass<FLUX, LIMC, R, TX, TY, IeeeOps, IeeeOps, GEOM, NT, SPLIT>(a, sm, i0, j0, hold);
  } else {
    edge_pass<FLUX, LIMC, R, TX, TY, IeeeOps, IeeeOps, GEOM, NT, SPLIT>(a, sm, i0, j0, hold);
  }
  __syncthreads();
  if constexpr (kHoldJ) {
    // the primitives are dead: their LDS takes the held j-fluxes for the cells below
    const int cy = tid / TX, cx = tid - cy * TX;
#pragma unroll
    for (int v = 0; v < 4; ++v) sm.sgj[v][cy][cx] = hold[v];
    __syncthreads();
  }

  // ---- 3. divergence + stage combination ---------------------------------
  double lmin = __builtin_inf();
  // calculate_convergence sums (last stage, when requested): squares of
  // (new - first stage) in [0..3] and of the first stage in [4..7]
  const bool conv = a.last_stage && a.conv_first[0] != nullptr;
  double csq[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) csq[v] = 0.0;
  const double EPSD = Lim<double>::eps;
  for (int c = tid; c < TX * TY; c += NT) {
    int cy = c / TX, cx = c - cy * TX;
    int gi = i0 + cx, gj = j0 + cy;
    if (gi > ni || gj > nj) continue;
    double dxv, dyv, lb, lr, lt, ll, vol, inv_vol;
    bool pow2;
    if (!GEOM && a.ugeom) {
      // uniform rectilinear grid: the same values, computed once on the host
      dxv = a.ug[0];
      dyv = a.ug[1];
      lb = a.ug[2];
      lr = a.ug[3];
      lt = a.ug[4];
      ll = a.ug[5];
      vol = a.ug[6];
      pow2 = a.ugeom_pow2 != 0;
      inv_vol = kExact ? a.ug[7] : rcp64(vol);
    } else {
      if constexpr (GEOM) {
        // quad_cell geometry of the cell: min_dx / min_dy, bottom / right / top / left lengths, volume
        const size_t gc = (size_t)(gi + nh - 1) + (size_t)(gj + nh - 1) * (size_t)a.quad_ncx;
        const size_t n = (size_t)a.quad_n;
        dxv = a.quad[Q_MDX * n + gc];
        dyv = a.quad[Q_MDY * n + gc];
        lb = a.quad[Q_LEN0 * n + gc];
        lr = a.quad[Q_LEN1 * n + gc];
        lt = a.quad[Q_LEN2 * n + gc];
        ll = a.quad[Q_LEN3 * n + gc];
        vol = a.quad[Q_VOL * n + gc];
      } else {
        dxv = a.dxi[gi + nh - 1];
        dyv = a.dyj[gj + nh - 1];
        double l1 = a.lxi[gi + nh - 1], l2 = a.lyj[gj + nh - 1];
        // edge-length equalisation (quad_cell.f90:110-112) over (l1, l2, l1, l2)
        double mx = fmax2(l1, l2), mn = fmin2(l1, l2);
        lb = l1, lr = l2, lt = l1, ll = l2;
        if (mx - mn < 2.0 * EPSD) lb = lr = lt = ll = mx;
        vol = dxv * dyv;
      }
      uint64_t vb = (uint64_t)__double_as_longlong(vol);
      const uint64_t ve = (vb >> 52) & 0x7FF;
      pow2 = (vb & 0x000FFFFFFFFFFFFFull) == 0 && ve > 1 && ve < 2046;
      // 1 / 2^k = 2^-k exactly: the sign of vol with the biased exponent 2046 - ve
      const double inv_pow2 = __longlong_as_double((long long)((vb & 0x8000000000000000ull) | ((2046 - ve) << 52)));
      inv_vol = kExact ? (pow2 ? inv_pow2 : 0.0) : rcp64(vol);
    }
    size_t off = (size_t)(gj - 1) * a.pitch + (size_t)(gi - 1);
    const size_t fs = (size_t)a.fstride;
    double fst[4] = {0.0, 0.0, 0.0, 0.0};  // issued before the flux balance to hide their latency
    if (conv) {
#pragma unroll
      for (int v = 0; v < 4; ++v) fst[v] = a.conv_first[0][off + v * fs];
    }
    // precision form, rectilinear cells (lr == ll, lt == lb): lengths over the volume
    const double sx = kExact ? 0.0 : lr * inv_vol, sy = kExact ? 0.0 : lb * inv_vol;
    double nv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double Fi = (double)sfi[v][cy][cx + 1], Fim = (double)sfi[v][cy][cx];
      double Gj, Gjm;
      if constexpr (kHoldJ) {
        Gj = cy + 1 < TY ? (double)sm.sgj[v][cy + 1][cx] : (double)sfj[v][0][cx];
        Gjm = (double)hold[v];
      } else {
        Gj = (double)sfj[v][cy + 1][cx];
        Gjm = (double)sfj[v][cy][cx];
      }
      double L;
      if constexpr (kExact) {
        L = divergence(Fi, Fim, Gj, Gjm, lb, lr, lt, ll, vol, pow2, inv_vol);
      } else {
        // precision form: the same FP64 flux balance without the reference's
        // max-normalisation and flushes (they only shape rounding at 1e-16); on
        // a rectilinear grid the right and left lengths are equal, as are the
        // top and bottom ones, so each pair takes one product (the difference of
        // two binary32 fluxes is exact in FP64)
        if constexpr (GEOM)
          L = -(((Fi * lr) - (Fim * ll)) + ((Gj * lt) - (Gjm * lb))) * inv_vol;
        else
          L = -(((Fi - Fim) * sx) + ((Gj - Gjm) * sy));
      }
      // time_derivative's energy source (mod_fluid.f90:500-509): a plain sum
      if (v == 3 && a.src_E) L = L + a.src_E[off];
      // the field operators' sum (its epsilon flush only shapes rounding at 1e-16,
      // which the precision form leaves out as in its flux balance)
      auto sum = [](double x, double y) { return kExact ? op_add(x, y) : x + y; };
      double r;
      if (a.mode == MODE_RESIDUAL) {
        r = L;
      } else if (a.mode == MODE_FIRST) {
        r = sum(a.uk[0][off + v * fs], L * (a.c * dt));
      } else if (a.mode == MODE_BASE) {
        r = sum(ld_cell<kExact>(&a.u0[0][off + v * fs]), L * (dt * a.c));
      } else {
        double t = sum(ld_cell<kExact>(&a.u0[0][off + v * fs]) * a.a, a.uk[0][off + v * fs] * a.b);
        r = sum(t, L * (a.c * dt));
      }
      nv[v] = r;
      if (conv) {
        double d = r - fst[v];
        csq[v] = csq[v] + d * d;
        csq[4 + v] = csq[4 + v] + fst[v] * fst[v];
      }
      st_cell<kExact>(&a.out[0][off + v * fs], r);
    }
    if (a.mode == MODE_RESIDUAL) continue;
    // precision form: the next stage recovers the primitives of this state
    // itself; only the last stage needs them here (checks and CFL)
    if (!kExact && !a.last_stage) continue;
    // primitives of the new state for the next stage (assign_fluid ->
    // calculate_derived_quantities, mod_fluid.f90:564-591) and its NaN check
    double cs;
    Prim<double> q = kExact ? cons_to_prim_vote(nv[0], nv[1], nv[2], nv[3], a.gamma, a.gamma - 1.0, &cs)
                            : cons_to_prim_fast(nv[0], nv[1], nv[2], nv[3], a.gamma, a.gamma - 1.0, 1, &cs);
    if (kExact) {
      st_cell<kExact>(&a.prim_out[0][off], q.rho);
      st_cell<kExact>(&a.prim_out[0][off + fs], q.u);
      st_cell<kExact>(&a.prim_out[0][off + 2 * fs], q.v);
      st_cell<kExact>(&a.prim_out[0][off + 3 * fs], q.p);
    }
    if (kExact && OUTFLOW_COL && gi == ni && a.bc[0] == BC_OUTFLOW) {
      Prim<double> g = outflow_ghost(q, a.gamma, a.bc_pressure);
      a.prim_out[0][off + 1] = g.rho;
      a.prim_out[0][off + fs + 1] = g.u;
      a.prim_out[0][off + 2 * fs + 1] = g.v;
      a.prim_out[0][off + 3 * fs + 1] = g.p;
    }
    if (q.rho != q.rho || q.u != q.u || q.v != q.v || q.p != q.p) err |= 1;
    if (a.last_stage) {
      if (q.rho < 0.0) err |= 2;
      if (q.p < 0.0) err |= 4;
      // get_timestep (mod_fluid.f90:554-558); min_dx/min_dy = |dx|, |dy| on rectilinear cells
      // (quad cells: the non-negative min_dx / min_dy themselves)
      double val = a.cfl / (((absd(q.u) + cs) / absd(dxv)) + ((absd(q.v) + cs) / absd(dyv)));
      lmin = (val < lmin) ? val : lmin;
    }
  }

  if (conv) {
    const int w = tid >> 6;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      double t = wave_sum(csq[v]);
      if ((tid & 63) == 0) red.s[w][v] = t;
    }
    __syncthreads();
    if (tid < 8) {
      double t = red.s[0][tid];
#pragma unroll
      for (int k = 1; k < NT / 64; ++k) t = t + red.s[k][tid];
      a.block_conv[(size_t)tid * (size_t)(a.tiles_x * a.tiles_y) + tile] = t;
    }
  }

  if (a.last_stage || a.block_err) {
    double wm = wave_min(lmin);
    int we = wave_or(err);
    int w = tid >> 6;
    if ((tid & 63) == 0) {
      red_min[w] = wm;
      red_err[w] = we;
    }
    __syncthreads();
    if (tid == 0) {
      double m = red_min[0];
      int e = red_err[0];
#pragma unroll
      for (int k = 1; k < NT / 64; ++k) {
        m = (red_min[k] < m) ? red_min[k] : m;
        e |= red_err[k];
      }
      if (a.block_min) a.block_min[tile] = m;
      if (a.block_err && e) a.block_err[tile] |= e;
    }
  }
}

// Single-solve stage launch: one workgroup per tile, XCD-aware tile order.
template <int FLUX, int LIMC, typename R, int TX, int TY, bool GEOM = false>
__global__ __launch_bounds__((stage_threads<R, LIMC, TX, TY>())) void stage_kernel(StageArgs a) {
  constexpr int NT = stage_threads<R, LIMC, TX, TY>();
  constexpr bool SPLIT = split_edge_rounds<R, TX, TY, NT>();
  __shared__ TileSmem<LIMC, R, TX, TY, NT, SPLIT> sm;
  __shared__ RedSmemT<NT / 64> red;
  if (a.loop && a.loop->done) return;
  const double dt = a.loop ? a.loop->dt : a.dt_value;
  stage_tile<FLUX, LIMC, R, TX, TY, false, GEOM, NT, SPLIT>(a, xcd_tile(blockIdx.x, gridDim.x), dt, sm, red);
}

}  // namespace cato
