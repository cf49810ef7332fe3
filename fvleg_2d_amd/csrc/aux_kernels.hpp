// Small kernels around the fused stage: initial primitive->conserved
// conversion, primitive recovery for download, CFL partials from stored
// primitives, the single-workgroup reduce/finalize that advances the
// device-resident time loop, and the per-operator kernels behind the
// reference's plugin entry points (boundary fill, edge reconstruction, edge
// flux) used for operator-level parity.
#pragma once

#include "stage_kernel.hpp"

namespace cato {

struct FieldGeom {
  int ni, nj, nh, pitch;
  int bc[4];
  double gamma, cfl;
  double gas[3];  // gas_constants(gamma)
  double bc_pressure;
  const double *dxi, *dyj;
  const double *quad;  // per-cell quad_cell geometry (QuadSlot arrays), null on rectilinear grids
  int quad_ncx;
  long long quad_n;
};

// Strided 2D copy on the device: dst(i, j) = src(i, j) for i < ni, j < nj with
// leading dimensions lds / ldd (host transfers stay contiguous 1D copies).
static __global__ void k_repack(const double *__restrict__ src, int lds, double *__restrict__ dst, int ldd, int ni) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= ni) return;
  dst[(size_t)j * ldd + i] = src[(size_t)j * lds + i];
}

// equation_of_state.f90:297-324 on the interior.
static __global__ void k_prim_to_cons(FieldGeom g, const double *__restrict__ rho, const double *__restrict__ u,
                               const double *__restrict__ v, const double *__restrict__ p, double *__restrict__ U0,
                               double *__restrict__ U1, double *__restrict__ U2, double *__restrict__ U3) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= g.ni) return;
  size_t k = (size_t)j * g.pitch + i;
  double gm1 = g.gamma - 1.0;
  double one_over = 1.0 / gm1;  // field / real is field * (1/real) (field_2d.f90:1169-1190)
  double r = rho[k], uu = u[k], vv = v[k], pp = p[k];
  U0[k] = r;
  U1[k] = uu * r;
  U2[k] = vv * r;
  double t1 = pp * one_over;
  double t2 = r * 0.5;
  double t5 = op_add(uu * uu, vv * vv);
  U3[k] = op_add(t1, t2 * t5);
}

// fluid_t%calculate_derived_quantities for download (u, v, p, cs).
static __global__ void k_derived(FieldGeom g, const double *__restrict__ U0, const double *__restrict__ U1,
                          const double *__restrict__ U2, const double *__restrict__ U3, double *__restrict__ u,
                          double *__restrict__ v, double *__restrict__ p, double *__restrict__ cs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= g.ni) return;
  size_t k = (size_t)j * g.pitch + i;
  double c;
  Prim<double> q = cons_to_prim(U0[k], U1[k], U2[k], U3[k], g.gamma, g.gamma - 1.0, &c);
  u[k] = q.u;
  v[k] = q.v;
  p[k] = q.p;
  cs[k] = c;
}

// Sound speed of stored primitives (equation_of_state.f90:250-251).
static __global__ void k_sound_speed(FieldGeom g, const double *__restrict__ rho, const double *__restrict__ p,
                              double *__restrict__ cs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= g.ni) return;
  size_t k = (size_t)j * g.pitch + i;
  cs[k] = sqrtr(op_div(p[k], rho[k]) * g.gamma);
}

// CFL partial minima over the interior (get_timestep, mod_fluid.f90:554-558).
// from_prims: u, v unfiltered and cs from (rho, p) as after fluid_t%initialize;
// otherwise the primitives are recovered from the conserved state.
static __global__ __launch_bounds__(kThreads) void k_cfl_partial(FieldGeom g, int from_prims, const double *__restrict__ a0,
                                                          const double *__restrict__ a1, const double *__restrict__ a2,
                                                          const double *__restrict__ a3, double *__restrict__ block_min,
                                                          int *__restrict__ block_err) {
  __shared__ double rm[kThreads / 64];
  __shared__ int re[kThreads / 64];
  size_t n = (size_t)g.ni * g.nj;
  double m = __builtin_inf();
  int err = 0;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < n; t += (size_t)gridDim.x * kThreads) {
    int j = (int)(t / g.ni), i = (int)(t - (size_t)j * g.ni);
    size_t k = (size_t)j * g.pitch + i;
    double u, v, cs, rho, p;
    if (from_prims) {
      rho = a0[k];
      u = a1[k];
      v = a2[k];
      p = a3[k];
      cs = sqrtr(op_div(p, rho) * g.gamma);
    } else {
      Prim<double> q = cons_to_prim(a0[k], a1[k], a2[k], a3[k], g.gamma, g.gamma - 1.0, &cs);
      rho = q.rho;
      u = q.u;
      v = q.v;
      p = q.p;
    }
    if (rho != rho || u != u || v != v || p != p) err |= 1;
    if (rho < 0.0) err |= 2;
    if (p < 0.0) err |= 4;
    double dx, dy;
    if (g.quad) {  // min_dx / min_dy of the cell (non-negative)
      const size_t gc = (size_t)(i + g.nh) + (size_t)(j + g.nh) * (size_t)g.quad_ncx;
      dx = g.quad[Q_MDX * (size_t)g.quad_n + gc];
      dy = g.quad[Q_MDY * (size_t)g.quad_n + gc];
    } else {
      dx = absd(g.dxi[i + g.nh]);
      dy = absd(g.dyj[j + g.nh]);
    }
    double val = g.cfl / (((absd(u) + cs) / dx) + ((absd(v) + cs) / dy));
    m = (val < m) ? val : m;
  }
  double wm = wave_min(m);
  int we = wave_or(err);
  if ((threadIdx.x & 63) == 0) {
    rm[threadIdx.x >> 6] = wm;
    re[threadIdx.x >> 6] = we;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = rm[0];
    int ee = re[0];
    for (int k = 1; k < kThreads / 64; ++k) {
      mm = (rm[k] < mm) ? rm[k] : mm;
      ee |= re[k];
    }
    block_min[blockIdx.x] = mm;
    block_err[blockIdx.x] = ee;
  }
}

// Error code of error.f90:6-9 from the OR of sanity flags.
__device__ inline int err_code(int flags) {
  if (flags & 1) return 3;  // NANS_FOUND
  if (flags & 2) return 1;  // NEG_DENSITY
  if (flags & 4) return 2;  // NEG_PRESSURE
  return 0;
}

// dt of the next step (main.f90:140-158): the startup period's dt, else a
// constant dt or the given first dt, else the CFL minimum m.
__device__ inline double chosen_dt(const LoopState *L, double m) {
  if (L->startup) return (L->time <= L->startup_period) ? L->startup_dt : m;
  if (L->constant_dt > 0.0) return L->constant_dt;
  if (L->iter0 + L->iteration == 0 && L->initial_dt > 0.0) return L->initial_dt;
  return m;
}

// With exact landing, no step carries the state's time past the next contour
// time or max_time.  The loop moves next_output_time on one step after the
// time reached it, so a time already on it aims at the one after.
__device__ inline double next_dt(const LoopState *L, double m) {
  double d = chosen_dt(L, m);
  if (L->exact) {
    double out = L->next_output_time;
    if (!(out - L->master_time > 0.0)) out = out + L->contour_dt;
    const double to_out = out - L->master_time;
    const double to_end = L->max_time - L->master_time;
    if (to_out > 0.0 && d > to_out) d = to_out;
    if (to_end > 0.0 && d > to_end) d = to_end;
  }
  return d;
}

// Time-loop bookkeeping after a reduction of CFL minimum m and sanity flags e.
// mode 1: initialise (first dt); mode 2: end of a step of main.f90:138-198
// (master%time += dt; contour cadence; dt clip of the driver time).
__device__ inline void loop_update(LoopState *L, double m, int e, int mode) {
  if (mode == 1) {
    L->dt = next_dt(L, m);
    if (!(L->time < L->max_time && L->iteration < L->max_iterations)) L->done = 1;
  } else if (mode == 2) {
    const double EPS = Lim<double>::eps;
    L->master_time = L->master_time + L->dt;
    double new_time = L->master_time;
    if (absd(L->time - L->next_output_time) < EPS) L->next_output_time = L->next_output_time + L->contour_dt;
    if (new_time > L->next_output_time) {
      double delta_t = absd(L->next_output_time - L->time);
      new_time = L->time + delta_t;
    }
    L->time = new_time;
    L->iteration = L->iteration + 1;
    int code = err_code(e);
    if (code) {
      L->error = code;
      L->done = 1;
    }
    L->dt = next_dt(L, m);
    if (!(L->time < L->max_time && L->iteration < L->max_iterations)) L->done = 1;
  }
}

// Reduce per-workgroup partials into out[0] = min dt, out[1] = flags; clears flags.
// mode 0: plain reduction into out; mode 1: initialise the loop (first dt);
// mode 2: end of a step of the main.f90 time loop.
// With cv.sums set it also forms the convergence norms of the step: the block
// sums of squares are added in a fixed order (threads, then waves), so the
// norms are the same on every run; NORM2's own order agrees to a few ulps.
static __global__ __launch_bounds__(1024) void k_reduce_finalize(int nb, double *__restrict__ block_min,
                                                          int *__restrict__ block_err, double *__restrict__ out,
                                                          LoopState *__restrict__ L, int mode, ConvOut cv) {
  __shared__ double sm[16];
  __shared__ int se[16];
  __shared__ double ss[16][8];
  if (mode == 2 && L->done) return;
  // One workgroup reads every partial: the loop is latency-bound, so each
  // thread issues kU partials (and their 8 convergence sums) before using any.
  constexpr int kU = 4;
  const bool conv = cv.sums != nullptr;
  double m = __builtin_inf();
  int e = 0;
  double t[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) t[v] = 0.0;
  for (int b0 = threadIdx.x; b0 < nb; b0 += kU * blockDim.x) {
    double x[kU];
    int f[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int b = b0 + k * (int)blockDim.x;
      x[k] = b < nb ? block_min[b] : __builtin_inf();
      f[k] = b < nb ? block_err[b] : 0;
    }
    if (conv) {
      double c[kU][8];
#pragma unroll
      for (int k = 0; k < kU; ++k) {
        const int b = b0 + k * (int)blockDim.x;
#pragma unroll
        for (int v = 0; v < 8; ++v) c[k][v] = b < nb ? cv.sums[(size_t)v * nb + b] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < kU; ++k)
#pragma unroll
        for (int v = 0; v < 8; ++v) t[v] = t[v] + c[k][v];
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int b = b0 + k * (int)blockDim.x;
      m = (x[k] < m) ? x[k] : m;
      e |= f[k];
      if (b < nb && f[k]) block_err[b] = 0;
    }
  }
  m = wave_min(m);
  e = wave_or(e);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) / 64;
  if ((threadIdx.x & 63) == 0) {
    sm[w] = m;
    se[w] = e;
  }
  if (conv) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      double y = t[v];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) y = y + __shfl_xor(y, o, 64);
      if ((threadIdx.x & 63) == 0) ss[w][v] = y;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 0; k < nw; ++k) {
    m = (sm[k] < m) ? sm[k] : m;
    e |= se[k];
  }
  double nrm[4] = {0.0, 0.0, 0.0, 0.0};
  if (conv) {
    double sum[8];
    for (int v = 0; v < 8; ++v) {
      sum[v] = ss[0][v];
      for (int k = 1; k < nw; ++k) sum[v] = sum[v] + ss[k][v];
    }
    // norm2(last - first) / norm2(first); the momenta divide by max(norm2(first), 1)
    for (int v = 0; v < 4; ++v) {
      double den = sqrt(sum[4 + v]);
      if (v == 1 || v == 2) den = den > 1.0 ? den : 1.0;
      nrm[v] = sqrt(sum[v]) / den;
      cv.norms[v] = nrm[v];
    }
  }
  if (out) {
    out[0] = m;
    out[1] = (double)e;
  }
  const double step_dt = (mode == 2) ? L->dt : 0.0;
  loop_update(L, m, e, mode);
  if (mode == 2 && conv && cv.hist) {
    // write_residual_history (mod_fluid.f90:1207-1223) row of the step just closed
    double *row = cv.hist + (size_t)((L->iteration - 1) % cv.hist_cap) * kHistRow;
    row[0] = (double)L->iteration;
    row[1] = L->master_time;
    row[2] = step_dt;
    for (int v = 0; v < 4; ++v) row[3 + v] = nrm[v];
  }
}

static __global__ void k_set_scalar(double *dst, double value) { *dst = value; }

// Primitives of a stored state as the precision form's stages recover them
// (cons_to_prim_fast), for the host-facing reads of P[0] in that form.
static __global__ void k_prims_fast(FieldGeom g, const double *__restrict__ U0, const double *__restrict__ U1,
                                    const double *__restrict__ U2, const double *__restrict__ U3,
                                    double *__restrict__ P0, double *__restrict__ P1, double *__restrict__ P2,
                                    double *__restrict__ P3) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= g.ni) return;
  size_t k = (size_t)j * g.pitch + i;
  double cs;
  Prim<double> q = cons_to_prim_fast(U0[k], U1[k], U2[k], U3[k], g.gamma, g.gamma - 1.0, 0, &cs);
  P0[k] = q.rho;
  P1[k] = q.u;
  P2[k] = q.v;
  P3[k] = q.p;
}

// +x outflow ghost column of a primitive set: column ni+1 of every row takes
// outflow_ghost of column ni (the stage kernel keeps it current for the states
// it produces; this fills it for primitives that arrive from elsewhere).
static __global__ void k_outflow_column(FieldGeom g, double *rho, double *u, double *v, double *p) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.nj) return;
  size_t k = (size_t)j * g.pitch + (size_t)(g.ni - 1);
  Prim<double> q = outflow_ghost(Prim<double>{rho[k], u[k], v[k], p[k]}, g.gamma, g.bc_pressure);
  rho[k + 1] = q.rho;
  u[k + 1] = q.u;
  v[k + 1] = q.v;
  p[k + 1] = q.p;
}

// ------------------------- +x pressure input boundary ----------------------
// pressure_input_bc.f90:232-385.  Its ghosts depend on the largest pressure of
// column ni, so they come from a one-workgroup pass over the column before each
// stage (the stage kernel then reads them from column ni+1 of its primitives).

// Where a pass takes its boundary time and desired state from.
struct PressureInput {
  const double *tab;     // n times, then n pressures, then n densities (null: the constants)
  int n;
  double p_const, rho_const;
  // mode 0: evaluate at t_at; mode 1: a step's stage at (t0 if first_multistage
  // else the fluid time at the step's start) + dt, with the fluid time and dt
  // from *loop when set (device-side loop) or t_at / dt_at
  int mode, first_multistage;
  double t_at, dt_at, t0;
  const LoopState *loop;
};

// finterp's linear_interp_1d%evaluate (linear_interpolation_module.f90:730-762,
// interval search dintrv 1189-1323): the interval [t_k, t_k+1] holding x, or
// the first / last interval outside the table, where the line extrapolates.
__device__ inline double interp_linear(const double *t, const double *f, int n, double x) {
  int k;
  if (x >= t[n - 1]) {
    k = n - 2;
  } else if (x < t[0]) {
    k = 0;
  } else {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      int mid = (lo + hi) / 2;
      if (x < t[mid]) hi = mid; else lo = mid;
    }
    k = lo;
  }
  double q1 = (x - t[k]) / (t[k + 1] - t[k]);
  double p1 = 1.0 - q1;
  return p1 * f[k] + q1 * f[k + 1];
}

__device__ inline void pressure_input_desired(const PressureInput &a, double &pd, double &rd) {
  if (!a.tab) {
    pd = a.p_const;
    rd = a.rho_const;
    return;
  }
  double t;
  if (a.mode == 0) {
    t = a.t_at;
  } else {
    double tn = a.loop ? a.loop->master_time : a.t_at;
    double dt = a.loop ? a.loop->dt : a.dt_at;
    t = (a.first_multistage ? a.t0 : tn) + dt;
  }
  pd = interp_linear(a.tab, a.tab + a.n, a.n, t);
  rd = interp_linear(a.tab, a.tab + 2 * a.n, a.n, t);
}

// Ghost state of one row from its domain cell (rho_d, u_d, p_d) (the domain v
// is taken as 0, pressure_input_bc.f90:312): outlet (u > 0) subsonic_outlet
// unless supersonic with the column's largest pressure at or above pd, then
// supersonic_outlet; inlet supersonic: (rd, u_d, 0, pd); else subsonic_inlet
// with normal (1, 0) and flow angle 0 (inlet_outlet.f90:26-191).  pd or rd
// below tiny (an error stop in the reference): NaN.
__device__ inline Prim<double> pressure_input_ghost(double rho_d, double u_d, double p_d, double pd, double rd,
                                                    double pmax, double gamma) {
  const double TINY = 2.2250738585072014e-308;
  const double nan = __builtin_nan("");
  if (!(pd >= TINY) || !(rd >= TINY)) return Prim<double>{nan, nan, 0.0, nan};
  const double cs = sqrt(gamma * p_d / rho_d);
  const double mach_u = u_d / cs;
  if (mach_u > 0.0) {
    if (mach_u <= 1.0 || pmax < pd) {
      double rho_b = rho_d + (pd - p_d) / (cs * cs);
      double u_b = u_d + (1.0 * (p_d - pd)) / (rho_d * cs);
      return Prim<double>{rho_b, u_b, 0.0, pd};
    }
    return Prim<double>{rho_d, u_d, 0.0, p_d};
  }
  if (absd(mach_u) > 1.0) return Prim<double>{rd, u_d, 0.0, pd};
  // subsonic_inlet: v_dot_n = u*(-1) + 0*(-0), the domain speed |u|, R- and
  // the boundary sound speed of the total-pressure inlet, isentropic p and rho
  const double gm1 = gamma - 1.0;
  double cos_theta = 0.0;
  double v_dot_n = u_d * (-1.0) + 0.0 * (-0.0);
  double v_tot_d = sqrt(u_d * u_d + 0.0 * 0.0);
  if (v_tot_d > 0.0) cos_theta = -(v_dot_n / v_tot_d);
  double cs_0_sq = cs * cs + (gm1 / 2.0) * (v_tot_d * v_tot_d);
  double R_minus = v_dot_n - (2.0 * cs / gm1);
  double den = gm1 * (cos_theta * cos_theta) + 2.0;
  double first_term = -(R_minus * gm1 / den);
  double cs_b = first_term;
  if (absd(cos_theta) > 0.0) {
    double second_term = (den * cs_0_sq) / (gm1 * (R_minus * R_minus));
    cs_b = first_term * (1.0 + cos_theta * sqrt(second_term - gm1 / 2.0));
  }
  double R_plus = v_dot_n + (2.0 * cs_b / gm1);
  double v_tot_b = (2.0 * cs_b / gm1) - R_plus;
  double mach_b = v_tot_b / cs_b;
  double base = 1.0 + 0.5 * gm1 * (mach_b * mach_b);
  return Prim<double>{rd * pow(base, -(1.0 / gm1)), v_tot_b * 1.0, 0.0, pd * pow(base, -(gamma / gm1))};
}

__device__ inline double nan_skipping_max(double m, double x) { return (x == x && !(m >= x)) ? x : m; }

// Largest pressure of a column over the workgroup (1024 threads); values come
// from get(j), j in [0, n).
template <typename Get> __device__ inline double column_max(int n, Get &&get) {
  __shared__ double red[16];
  const int tid = threadIdx.x;
  double m = __builtin_nan("");
  for (int j = tid; j < n; j += blockDim.x) m = nan_skipping_max(m, get(j));
  for (int o = 32; o > 0; o >>= 1) m = nan_skipping_max(m, __shfl_xor(m, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = nan_skipping_max(r, red[w]);
  return r;
}

// Column ni+1 of a pitched primitive set from column ni: of the primitives
// themselves (src = the same set) or, RECOVER, of the state src = U_k
// recovered as the precision form's stages recover it.  One workgroup.
template <bool RECOVER>
__device__ inline void pressure_column(const FieldGeom &g, const PressureInput &a, const double *s0,
                                       const double *s1, const double *s2, const double *s3, double *rho, double *u,
                                       double *v, double *p) {
  if (a.loop && a.loop->done) return;
  auto cell = [&](int j) {
    size_t k = (size_t)j * g.pitch + (size_t)(g.ni - 1);
    if constexpr (RECOVER) {
      double cs;
      return cons_to_prim_fast(s0[k], s1[k], s2[k], s3[k], g.gamma, g.gamma - 1.0, 0, &cs);
    } else {
      return Prim<double>{s0[k], s1[k], s2[k], s3[k]};
    }
  };
  const double pmax = column_max(g.nj, [&](int j) { return cell(j).p; });
  double pd, rd;
  pressure_input_desired(a, pd, rd);
  for (int j = threadIdx.x; j < g.nj; j += blockDim.x) {
    Prim<double> d = cell(j);
    Prim<double> q = pressure_input_ghost(d.rho, d.u, d.p, pd, rd, pmax, g.gamma);
    size_t k = (size_t)j * g.pitch + (size_t)g.ni;
    rho[k] = q.rho;
    u[k] = q.u;
    v[k] = q.v;
    p[k] = q.p;
  }
}

template <bool RECOVER>
static __global__ __launch_bounds__(1024) void k_pressure_column(FieldGeom g, PressureInput a, const double *s0,
                                                                 const double *s1, const double *s2,
                                                                 const double *s3, double *rho, double *u, double *v,
                                                                 double *p) {
  pressure_column<RECOVER>(g, a, s0, s1, s2, s3, rho, u, v, p);
}

// The pass of one stage for several solves of a batch or a server (one
// workgroup each): where it reads (the stage's primitives, or its U_k to
// recover) and the primitive set whose column ni+1 it writes, four fields
// `fs` elements apart.
struct PressurePass {
  FieldGeom g;
  PressureInput pa;
  const double *src;
  double *dst;
  long long fs;
  int recover;
};

static __global__ __launch_bounds__(1024) void k_pressure_columns(const PressurePass *passes) {
  const PressurePass &q = passes[blockIdx.x];
  const size_t fs = (size_t)q.fs;
  if (q.recover)
    pressure_column<true>(q.g, q.pa, q.src, q.src + fs, q.src + 2 * fs, q.src + 3 * fs, q.dst, q.dst + fs,
                          q.dst + 2 * fs, q.dst + 3 * fs);
  else
    pressure_column<false>(q.g, q.pa, q.src, q.src + fs, q.src + 2 * fs, q.src + 3 * fs, q.dst, q.dst + fs,
                           q.dst + 2 * fs, q.dst + 3 * fs);
}

// The same on halo'd arrays (leading dimension ld) after the other sides are
// filled: every row of the halo'd extent, every ghost layer (apply order of
// flux_solver.f90:157-183, the pressure input last at priority 0).
static __global__ __launch_bounds__(1024) void k_pressure_halo(FieldGeom g, PressureInput a, int ld, double *rho,
                                                               double *u, double *v, double *p) {
  const int ny = g.nj + 2 * g.nh, ii = g.ni + g.nh - 1;
  const double pmax = column_max(ny, [&](int jj) { return p[(size_t)ii + (size_t)jj * ld]; });
  double pd, rd;
  pressure_input_desired(a, pd, rd);
  for (int jj = threadIdx.x; jj < ny; jj += blockDim.x) {
    size_t k = (size_t)ii + (size_t)jj * ld;
    Prim<double> q = pressure_input_ghost(rho[k], u[k], p[k], pd, rd, pmax, g.gamma);
    for (int l = 1; l <= g.nh; ++l) {
      rho[k + l] = q.rho;
      u[k + l] = q.u;
      v[k + l] = q.v;
      p[k + l] = q.p;
    }
  }
}

// --------------------------- per-operator kernels --------------------------
// Halo'd column-major arrays: element (i, j), i,j in [1-nh, n+nh], at
// (i+nh-1) + (j+nh-1)*ld.

// apply_primitive_bc: every ghost cell takes the value of its mapped interior cell.
static __global__ void k_apply_bc(FieldGeom g, int ld, double *rho, double *u, double *v, double *p) {
  int nx = g.ni + 2 * g.nh, ny = g.nj + 2 * g.nh;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny) return;
  int ii = t % nx, jj = t / nx;
  int gi = ii + 1 - g.nh, gj = jj + 1 - g.nh;
  if (gi >= 1 && gi <= g.ni && gj >= 1 && gj <= g.nj) return;
  bool fu, fv;
  int si = map_axis(gi, g.ni, g.bc[0], g.bc[1], fu);
  int sj = map_axis(gj, g.nj, g.bc[2], g.bc[3], fv);
  size_t d = (size_t)ii + (size_t)jj * ld;
  size_t s = (size_t)(si + g.nh - 1) + (size_t)(sj + g.nh - 1) * ld;
  Prim<double> q{rho[s], fu ? -u[s] : u[s], fv ? -v[s] : v[s], p[s]};
  if (gi > g.ni && g.bc[0] == BC_OUTFLOW) q = outflow_ghost(q, g.gamma, g.bc_pressure);
  rho[d] = q.rho;
  u[d] = q.u;
  v[d] = q.v;
  p[d] = q.p;
}

// muscl_interpolation_t%interpolate_edge_values on a halo'd field q:
// edges (2, 0:ni, 0:nj) for both directions.
template <typename R, int LIMC>
static __global__ void k_interpolate(FieldGeom g, int ld, const double *__restrict__ q, double *__restrict__ ie,
                              double *__restrict__ je) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int n = (g.ni + 1) * (g.nj + 1);
  if (t >= 2 * n) return;
  int dir = t >= n;
  int k = t - dir * n;
  int i = k % (g.ni + 1), j = k / (g.ni + 1);
  // local copy of the stencil neighbourhood as R
  R buf[7][3];
  for (int b = -1; b <= 1; ++b)
    for (int s = -3; s <= 3; ++s) {
      int gi = dir == 0 ? i + s : i + b;
      int gj = dir == 0 ? j + b : j + s;
      R val = R(0);
      if (gi >= 1 - g.nh && gi <= g.ni + g.nh && gj >= 1 - g.nh && gj <= g.nj + g.nh)
        val = (R)q[(size_t)(gi + g.nh - 1) + (size_t)(gj + g.nh - 1) * ld];
      buf[s + 3][b + 1] = val;
    }
  R L, Rv;
  // normal stride 3 (rows of buf), transverse stride 1
  recon_var<R, LIMC>(&buf[3][1], 3, 1, L, Rv);
  double *dst = dir == 0 ? ie : je;
  dst[2 * (size_t)k + 0] = (double)L;
  dst[2 * (size_t)k + 1] = (double)Rv;
}

// get_<dir>flux on given edge states; p is the halo'd cell pressure.
template <typename R, int FLUX>
static __global__ void k_edge_flux(FieldGeom g, int ld, const double *__restrict__ edges, const double *__restrict__ p,
                            int low_mach_fix, double *__restrict__ iflux, double *__restrict__ jflux) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nI = (g.ni + 1) * g.nj, nJ = g.ni * (g.nj + 1);
  if (t >= nI + nJ) return;
  int dir = t >= nI;
  int i, j;
  if (!dir) {
    i = t % (g.ni + 1);
    j = t / (g.ni + 1) + 1;
  } else {
    int k = t - nI;
    i = k % g.ni + 1;
    j = k / g.ni;
  }
  size_t ne = (size_t)2 * (g.ni + 1) * (g.nj + 1);
  size_t e = 2 * ((size_t)i + (size_t)(g.ni + 1) * j);
  const double *b = edges + (dir ? ne : 0);
  EdgeState<R> s;
  s.rL = (R)b[0 * 2 * ne + e]; s.rR = (R)b[0 * 2 * ne + e + 1];
  s.uL = (R)b[1 * 2 * ne + e]; s.uR = (R)b[1 * 2 * ne + e + 1];
  s.vL = (R)b[2 * 2 * ne + e]; s.vR = (R)b[2 * 2 * ne + e + 1];
  s.pL = (R)b[3 * 2 * ne + e]; s.pR = (R)b[3 * 2 * ne + e + 1];
  auto P = [&](int a, int c) -> R {
    int gi = dir == 0 ? i + a : i + c;
    int gj = dir == 0 ? j + c : j + a;
    return (R)p[(size_t)(gi + g.nh - 1) + (size_t)(gj + g.nh - 1) * ld];
  };
  EdgeState<R> sb = s;
  if constexpr (needs_superbee_states<FLUX>()) {  // superbee states follow the 8 arrays
    const double *q = b + 8 * ne;
    sb.rL = (R)q[0 * 2 * ne + e]; sb.rR = (R)q[0 * 2 * ne + e + 1];
    sb.uL = (R)q[1 * 2 * ne + e]; sb.uR = (R)q[1 * 2 * ne + e + 1];
    sb.vL = (R)q[2 * 2 * ne + e]; sb.vR = (R)q[2 * 2 * ne + e + 1];
    sb.pL = (R)q[3 * 2 * ne + e]; sb.pR = (R)q[3 * 2 * ne + e + 1];
  }
  PressNbr<R> pn;
  pn.p00 = P(0, 0); pn.p10 = P(1, 0); pn.p20 = P(2, 0); pn.pm10 = P(-1, 0);
  pn.p0p = P(0, 1); pn.p0m = P(0, -1); pn.p1p = P(1, 1); pn.p1m = P(1, -1);
  R f[4];
  const Gas<R> gas = gas_of<R>(g.gamma, g.gas);
  if (g.quad) {  // normal of the right (i-edge) or top (j-edge) edge of cell (i, j)
    const size_t gc = (size_t)(i + g.nh - 1) + (size_t)(j + g.nh - 1) * (size_t)g.quad_ncx;
    const size_t n = (size_t)g.quad_n;
    s.nx = (R)g.quad[(dir ? Q_NTX : Q_NRX) * n + gc];
    s.ny = (R)g.quad[(dir ? Q_NTY : Q_NRY) * n + gc];
    if (!dir)
      edge_flux<R, FLUX, 2>(gas, low_mach_fix, s, pn, f, &sb);
    else
      edge_flux<R, FLUX, 3>(gas, low_mach_fix, s, pn, f, &sb);
  } else if (!dir) {
    edge_flux<R, FLUX, 0>(gas, low_mach_fix, s, pn, f, &sb);
  } else {
    edge_flux<R, FLUX, 1>(gas, low_mach_fix, s, pn, f, &sb);
  }
  double *d = dir ? jflux + 4 * ((size_t)(i - 1) + (size_t)g.ni * j) : iflux + 4 * ((size_t)i + (size_t)(g.ni + 1) * (j - 1));
  for (int v = 0; v < 4; ++v) d[v] = (double)f[v];
}

}  // namespace cato
