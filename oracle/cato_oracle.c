/*
 * CATO hot-path oracle (test infrastructure only; see cato_oracle.h).
 *
 * Every function restates the reference loop it cites, in the reference's
 * operation order, so that a build with -ffp-contract=off and IEEE division
 * and sqrt reproduces the reference arithmetic exactly.  Fortran semantics
 * used throughout:
 *   - x**2 is x*x; a*b*c is (a*b)*c; a+b+c is (a+b)+c (left to right).
 *   - MAX(a1,a2,...) folds left keeping the running value unless a later
 *     argument compares greater (MIN: less); SIGN(1,x) is copysign(1,x).
 *   - epsilon(1.0_rk) = 2^-52, tiny(1.0_rk) = DBL_MIN.
 */
#include "cato_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS DBL_EPSILON
#define TINY DBL_MIN

/* ORACLE_OMP (liboracle_omp.so, CPU-baseline timing only): rows of the
 * cell/edge loops in parallel; every cell is computed exactly as in the serial
 * build, so the results are the same. */
#ifdef ORACLE_OMP
#define OR_PAR _Pragma("omp parallel for schedule(static)")
#else
#define OR_PAR
#endif

static inline double fmax2(double a, double b) { return (b > a) ? b : a; }
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }
static inline double fmax3(double a, double b, double c) { return fmax2(fmax2(a, b), c); }
static inline double fmin3(double a, double b, double c) { return fmin2(fmin2(a, b), c); }
static inline double sq(double x) { return x * x; }

int or_ncx(int ni, int nh) { return ni + 2 * nh; }
int or_ncy(int nj, int nh) { return nj + 2 * nh; }

/* column-major halo'd cell index, 1-based cell coordinates */
#define CI(i, j) ((size_t)((i) + nh - 1) + (size_t)((j) + nh - 1) * (size_t)(ni + 2 * nh))
/* edge state arrays (2, 0:ni, 0:nj) */
#define EI(k, i, j) ((size_t)(k) + 2 * ((size_t)(i) + (size_t)(ni + 1) * (size_t)(j)))
/* iflux (4, 0:ni, 1:nj) and jflux (4, 1:ni, 0:nj) */
#define FI(v, i, j) ((size_t)(v) + 4 * ((size_t)(i) + (size_t)(ni + 1) * (size_t)((j) - 1)))
#define FJ(v, i, j) ((size_t)(v) + 4 * ((size_t)((i) - 1) + (size_t)ni * (size_t)(j)))

/* ------------------------------------------------------------------------- */
/* Geometry: quad_cell.f90:69-229, grid_block_2d.f90:352-391                   */
/* ------------------------------------------------------------------------- */
int or_geometry_from_nodes(int ni, int nj, int nh, const double *node_x, const double *node_y, or_geometry *g) {
  const int nnx = ni + 2 * nh + 1;
#define NI(i, j) ((size_t)((i) + nh - 1) + (size_t)((j) + nh - 1) * (size_t)nnx)
  g->ni = ni;
  g->nj = nj;
  g->nh = nh;
  int bad = 0;
  for (int j = 1 - nh; j <= nj + nh; ++j) {
    for (int i = 1 - nh; i <= ni + nh; ++i) {
      double x[4] = {node_x[NI(i, j)], node_x[NI(i + 1, j)], node_x[NI(i + 1, j + 1)], node_x[NI(i, j + 1)]};
      double y[4] = {node_y[NI(i, j)], node_y[NI(i + 1, j)], node_y[NI(i + 1, j + 1)], node_y[NI(i, j + 1)]};
      /* round-off checks (quad_cell.f90:78-91) */
      if (fabs(y[1] - y[0]) < EPS) y[1] = y[0];
      if (fabs(y[3] - y[2]) < EPS) y[3] = y[2];
      if (fabs(x[2] - x[1]) < EPS) x[2] = x[1];
      if (fabs(x[3] - x[0]) < EPS) x[3] = x[0];
      /* volume (quad_cell.f90:160-175) */
      double vol = (x[1] - x[0]) * (y[3] - y[0]) - (x[3] - x[0]) * (y[1] - y[0]);
      if (vol <= 0.0) bad = 1;
      /* edge lengths + equalization (quad_cell.f90:100-112) */
      double el[4];
      el[0] = sqrt(sq(x[1] - x[0]) + sq(y[1] - y[0]));
      el[1] = sqrt(sq(x[2] - x[1]) + sq(y[2] - y[1]));
      el[2] = sqrt(sq(x[3] - x[2]) + sq(y[3] - y[2]));
      el[3] = sqrt(sq(x[0] - x[3]) + sq(y[0] - y[3]));
      double mx = el[0], mn = el[0];
      for (int e = 1; e < 4; ++e) {
        mx = fmax2(mx, el[e]);
        mn = fmin2(mn, el[e]);
      }
      if (mx - mn < 2.0 * EPS) el[0] = el[1] = el[2] = el[3] = mx;
      /* midpoints (quad_cell.f90:114-117) */
      double mxp[4], myp[4];
      mxp[0] = (x[1] + x[0]) / 2.0; myp[0] = (y[1] + y[0]) / 2.0;
      mxp[1] = (x[2] + x[1]) / 2.0; myp[1] = (y[2] + y[1]) / 2.0;
      mxp[2] = (x[3] + x[2]) / 2.0; myp[2] = (y[3] + y[2]) / 2.0;
      mxp[3] = (x[0] + x[3]) / 2.0; myp[3] = (y[0] + y[3]) / 2.0;
      /* min dx / dy (quad_cell.f90:119-123) */
      double mdx = fmin2(fabs(x[1] - x[0]), fabs(x[2] - x[3]));
      double mdy = fmin2(fabs(y[3] - y[0]), fabs(y[2] - y[1]));
      /* normals (quad_cell.f90:193-229, vector.f90:58-96) */
      static const int head[4] = {1, 2, 3, 0};
      size_t c = CI(i, j);
      for (int e = 0; e < 4; ++e) {
        double ddx = x[head[e]] - mxp[e];
        double ddy = y[head[e]] - myp[e];
        double vx = (mxp[e] + ddy) - mxp[e];
        double vy = (myp[e] - ddx) - myp[e];
        double len = sqrt(sq(vx) + sq(vy));
        g->edge_norm[8 * c + 2 * e + 0] = vx / len;
        g->edge_norm[8 * c + 2 * e + 1] = vy / len;
        g->edge_lengths[4 * c + e] = el[e];
      }
      g->volume[c] = vol;
      g->dx[c] = mdx;
      g->dy[c] = mdy;
    }
  }
#undef NI
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* EOS / field-operator arithmetic (field_2d_cpu_operators.f90:84-423)        */
/* ------------------------------------------------------------------------- */
static inline double op_add(double a, double b) { double r = a + b; return (fabs(r) < EPS) ? 0.0 : r; }
static inline double op_sub(double a, double b) { double r = a - b; return (fabs(r) < EPS) ? 0.0 : r; }
static inline double op_div(double a, double b) { return (fabs(a) < TINY) ? 0.0 : a / b; }

void or_prim_to_cons(const or_config *c, const double *rho, const double *u, const double *v, const double *p,
                     double *rho_u, double *rho_v, double *rho_E) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const double gm1 = c->gamma - 1.0;
  const double one_over = 1.0 / gm1; /* field_div_real_1d multiplies by 1/x (field_2d.f90:1169-1190) */
  OR_PAR
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i) {
      size_t k = CI(i, j);
      rho_u[k] = u[k] * rho[k];
      rho_v[k] = v[k] * rho[k];
      double t1 = p[k] * one_over;
      double t2 = rho[k] * 0.5;
      double t5 = op_add(u[k] * u[k], v[k] * v[k]);
      rho_E[k] = op_add(t1, t2 * t5);
    }
}

void or_sound_speed(const or_config *c, const double *rho, const double *p, double *cs) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  OR_PAR
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i) {
      size_t k = CI(i, j);
      cs[k] = sqrt(op_div(p[k], rho[k]) * c->gamma);
    }
}

void or_derived_quantities(const or_config *c, const double *rho, const double *rho_u, const double *rho_v,
                           const double *rho_E, double *u, double *v, double *p, double *cs) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const double gm1 = c->gamma - 1.0;
  OR_PAR
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i) {
      size_t k = CI(i, j);
      /* equation_of_state.f90:286-292 */
      double uu = op_div(rho_u[k], rho[k]);
      double vv = op_div(rho_v[k], rho[k]);
      double t1 = rho[k] * gm1;
      double t2 = op_div(rho_E[k], rho[k]);
      double t5 = op_add(uu * uu, vv * vv);
      double t6 = t5 * 0.5;
      double t7 = op_sub(t2, t6);
      double pp = t1 * t7;
      /* equation_of_state.f90:250-251 */
      double c2 = op_div(pp, rho[k]) * c->gamma;
      double ss = sqrt(c2);
      /* mod_fluid.f90:576-588 */
      double mu = op_div(uu, ss);
      double mv = op_div(vv, ss);
      if (fabs(mv) < 1e-13) vv = 0.0;
      if (fabs(mu) < 1e-13) uu = 0.0;
      u[k] = uu;
      v[k] = vv;
      p[k] = pp;
      cs[k] = ss;
    }
}

/* ------------------------------------------------------------------------- */
/* Boundary conditions on one image                                           */
/* ------------------------------------------------------------------------- */
/* periodic_bc.fypp:65, zero_gradient_bc.f90:50, symmetry_bc.f90:52, outlet_bc.f90:54 */
/* pressure_input_bc.f90:78 */
static int bc_priority(int t) { return t == OR_BC_PERIODIC ? 2 : (t == OR_BC_OUTFLOW || t == OR_BC_PRESSURE_INPUT) ? 0 : 1; }

/* periodic_bc.fypp:78-283: either x side fills both x ghost strips (rows jlo:jhi);
 * corners only when all four sides are periodic. */
static void bc_periodic(const or_config *c, int xdir, double *f) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  int do_corners = c->bc_plus_x == OR_BC_PERIODIC && c->bc_minus_x == OR_BC_PERIODIC &&
                   c->bc_plus_y == OR_BC_PERIODIC && c->bc_minus_y == OR_BC_PERIODIC;
  if (xdir) {
    for (int j = 1; j <= nj; ++j)
      for (int g = 0; g < nh; ++g) {
        f[CI(ni + 1 + g, j)] = f[CI(1 + g, j)];
        f[CI(1 - nh + g, j)] = f[CI(ni - nh + 1 + g, j)];
      }
  } else {
    for (int g = 0; g < nh; ++g)
      for (int i = 1; i <= ni; ++i) {
        f[CI(i, 1 - nh + g)] = f[CI(i, nj - nh + 1 + g)];
        f[CI(i, nj + 1 + g)] = f[CI(i, 1 + g)];
      }
  }
  if (do_corners) {
    for (int b = 0; b < nh; ++b)
      for (int a = 0; a < nh; ++a) {
        f[CI(1 - nh + a, 1 - nh + b)] = f[CI(ni - nh + 1 + a, nj - nh + 1 + b)];
        f[CI(ni + 1 + a, 1 - nh + b)] = f[CI(1 + a, nj - nh + 1 + b)];
        f[CI(1 - nh + a, nj + 1 + b)] = f[CI(ni - nh + 1 + a, 1 + b)];
        f[CI(ni + 1 + a, nj + 1 + b)] = f[CI(1 + a, 1 + b)];
      }
  }
}

/* zero_gradient_bc.f90:54-121 and symmetry_bc.f90:56-133: whole transverse extent. */
static void bc_copy(const or_config *c, int loc, int mirror, int negate, double *f) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  for (int g = 1; g <= nh; ++g) {
    switch (loc) {
      case 0: /* +x */
        for (int j = 1 - nh; j <= nj + nh; ++j) {
          double s = f[CI(mirror ? ni - (g - 1) : ni, j)];
          f[CI(ni + g, j)] = negate ? -s : s;
        }
        break;
      case 1: /* -x */
        for (int j = 1 - nh; j <= nj + nh; ++j) {
          double s = f[CI(mirror ? 1 + (g - 1) : 1, j)];
          f[CI(1 - g, j)] = negate ? -s : s;
        }
        break;
      case 2: /* +y */
        for (int i = 1 - nh; i <= ni + nh; ++i) {
          double s = f[CI(i, mirror ? nj - (g - 1) : nj)];
          f[CI(i, nj + g)] = negate ? -s : s;
        }
        break;
      default: /* -y */
        for (int i = 1 - nh; i <= ni + nh; ++i) {
          double s = f[CI(i, mirror ? 1 + (g - 1) : 1)];
          f[CI(i, 1 - g)] = negate ? -s : s;
        }
        break;
    }
  }
}

/* outlet_bc.f90:96-123: at +x every ghost layer of row j (j over the whole halo'd extent) takes the
 * last interior cell of the row, extrapolated along the outgoing characteristic to p_b when that cell
 * is subsonic.  The reference skips a row whose rho or p_b is not > tiny (its ghost keeps an older
 * value); this restatement copies the cell there instead, as the HIP path does. */
static void bc_outflow_plus_x(const or_config *c, double *rho, double *u, double *v, double *p) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const double gamma = c->gamma, p_b = c->bc_pressure; /* TINY = tiny(1.0_rk) */
  for (int i = ni + 1; i <= ni + nh; ++i)
    for (int j = 1 - nh; j <= nj + nh; ++j) {
      const size_t s = CI(ni, j), d = CI(i, j);
      const double p_d = p[s], rho_0 = rho[s], u_d = u[s], v_d = v[s];
      double gr = rho_0, gu = u_d, gv = v_d, gp = p_d;
      if (p_b > TINY && rho_0 > TINY) {
        double cs = sqrt(gamma * fabs(p_d / rho_0));
        double vel = sqrt(u_d * u_d + v_d * v_d);
        double mach = vel / cs;
        double nx = u_d / (vel + 1e-30), ny = v_d / (vel + 1e-30);
        if (!(mach > 1.0)) {
          gr = rho_0 + (p_b - p_d) / (cs * cs);
          gu = u_d + nx * (p_d - p_b) / (rho_0 * cs);
          gv = v_d + ny * (p_d - p_b) / (rho_0 * cs);
          gp = p_b;
        }
      }
      rho[d] = gr;
      u[d] = gu;
      v[d] = gv;
      p[d] = gp;
    }
}

/* linear_interp_1d%evaluate: interp_1d (linear_interpolation_module.f90:730-762) on the interval dintrv
 * (1189-1323) returns for strictly increasing knots, [t_k, t_k+1] with t_k <= x < t_k+1, and the first or
 * last interval outside the table, where the same formula extrapolates. */
double or_interp_linear(int n, const double *t, const double *f, double x) {
  int k;
  if (x >= t[n - 1]) {
    k = n - 2;
  } else if (x < t[0]) {
    k = 0;
  } else {
    int lo = 0, hi = n - 1; /* t[lo] <= x < t[hi] */
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

/* subsonic_inlet (inlet_outlet.f90:26-140) for the +x boundary: normal (1, 0), flow angle 0, and the
 * domain v that pressure_input_bc passes (0).  Operation order as written there; x**2 is x*x and
 * x**y is pow (what gfortran emits). */
static void subsonic_inlet_px(double gamma, double rho, double u, double p, double rho_0, double p_0,
                              double *rho_b, double *u_b, double *p_b) {
  const double n_x = 1.0, n_y = 0.0, v = 0.0, theta = 0.0;
  double cos_theta = 0.0;
  double v_dot_n = u * (-n_x) + v * (-n_y);
  double v_tot_d = sqrt(u * u + v * v);
  if (v_tot_d > 0.0) cos_theta = -(v_dot_n / v_tot_d);
  double cs_d = sqrt(gamma * p / rho);
  double cs_0_sq = cs_d * cs_d + ((gamma - 1.0) / 2.0) * (v_tot_d * v_tot_d);
  double R_minus = v_dot_n - (2.0 * cs_d / (gamma - 1.0));
  double first_term = -(R_minus * (gamma - 1.0) / ((gamma - 1.0) * (cos_theta * cos_theta) + 2.0));
  double cs_b;
  if (fabs(cos_theta) > 0.0) {
    double second_term =
        (((gamma - 1.0) * (cos_theta * cos_theta) + 2.0) * cs_0_sq) / ((gamma - 1.0) * (R_minus * R_minus));
    double third_term = (gamma - 1.0) / 2.0;
    cs_b = first_term * (1.0 + cos_theta * sqrt(second_term - third_term));
  } else {
    cs_b = first_term;
  }
  double R_plus = v_dot_n + (2.0 * cs_b / (gamma - 1.0));
  double v_tot_b = (2.0 * cs_b / (gamma - 1.0)) - R_plus;
  double mach_b = v_tot_b / cs_b;
  double base = 1.0 + 0.5 * (gamma - 1.0) * (mach_b * mach_b);
  *u_b = v_tot_b * cos(theta);
  *p_b = p_0 * pow(base, -(gamma / (gamma - 1.0)));
  *rho_b = rho_0 * pow(base, -(1.0 / (gamma - 1.0)));
}

/* pressure_input_bc_t%apply at +x (pressure_input_bc.f90:232-385): the desired boundary pressure and
 * density at bc_time (get_desired_pressure / get_desired_density, 181-230: the constant inputs, or the
 * table), then for every row j of the halo'd extent the state of mod_inlet_outlet chosen by the domain
 * Mach number u/cs of cell (ni, j) and, for an outlet, by the largest pressure of column ni; every ghost
 * layer of the row takes it with v = 0.  The reference stops with an error when the desired pressure or
 * density is below tiny; this restatement fills the ghosts with NaN instead, which the step reports as
 * NaN (error code 3), as the HIP path does.  maxval skips NaNs as gfortran's does. */
static void bc_pressure_input_plus_x(const or_config *c, double *rho, double *u, double *v, double *p) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const double gamma = c->gamma;
  double pd, rd;
  if (c->pi_n >= 2) {
    pd = or_interp_linear(c->pi_n, c->pi_t, c->pi_p, c->bc_time);
    rd = or_interp_linear(c->pi_n, c->pi_t, c->pi_rho, c->bc_time);
  } else {
    pd = c->bc_pressure;
    rd = c->bc_density;
  }
  const int bad = !(pd >= TINY) || !(rd >= TINY);
  double pmax = NAN;
  for (int j = 1 - nh; j <= nj + nh; ++j) {
    double x = p[CI(ni, j)];
    if (x == x && !(pmax >= x)) pmax = x;
  }
  for (int j = 1 - nh; j <= nj + nh; ++j) {
    const double rho_d = rho[CI(ni, j)], u_d = u[CI(ni, j)], p_d = p[CI(ni, j)];
    double gr, gu, gp;
    if (bad) {
      gr = gu = gp = NAN;
    } else {
      double cs = sqrt(gamma * p_d / rho_d);
      double mach_u = u_d / cs;
      if (mach_u > 0.0) {
        if (mach_u <= 1.0 || pmax < pd) {
          /* subsonic_outlet (inlet_outlet.f90:142-177), exit pressure pd, normal (1, 0) */
          double cs_d = sqrt(gamma * p_d / rho_d);
          gr = rho_d + (pd - p_d) / (cs_d * cs_d);
          gu = u_d + (1.0 * (p_d - pd)) / (rho_d * cs_d);
          gp = pd;
        } else { /* supersonic_outlet (186-191) */
          gr = rho_d;
          gu = u_d;
          gp = p_d;
        }
      } else if (fabs(mach_u) > 1.0) { /* pressure_input_bc.f90:335-336 */
        gr = rd;
        gu = u_d;
        gp = pd;
      } else {
        subsonic_inlet_px(gamma, rho_d, u_d, p_d, rd, pd, &gr, &gu, &gp);
      }
    }
    for (int i = ni + 1; i <= ni + nh; ++i) {
      rho[CI(i, j)] = gr;
      u[CI(i, j)] = gu;
      v[CI(i, j)] = 0.0;
      p[CI(i, j)] = gp;
    }
  }
}

static void bc_apply_one(const or_config *c, int type, int loc, double *rho, double *u, double *v, double *p) {
  int xdir = (loc == 0 || loc == 1);
  if (type == OR_BC_PRESSURE_INPUT) {
    /* pressure_input_bc.f90:377-381: +x only (the host rejects the other sides) */
    if (loc == 0) bc_pressure_input_plus_x(c, rho, u, v, p);
    return;
  }
  if (type == OR_BC_OUTFLOW) {
    if (loc == 0) {
      bc_outflow_plus_x(c, rho, u, v, p);
    } else { /* outlet_bc.f90:124-162: the other sides copy the last interior line over the full extent */
      bc_copy(c, loc, 0, 0, rho);
      bc_copy(c, loc, 0, 0, u);
      bc_copy(c, loc, 0, 0, v);
      bc_copy(c, loc, 0, 0, p);
    }
    return;
  }
  if (type == OR_BC_PERIODIC) {
    bc_periodic(c, xdir, rho);
    bc_periodic(c, xdir, u);
    bc_periodic(c, xdir, v);
    bc_periodic(c, xdir, p);
  } else if (type == OR_BC_ZERO_GRADIENT) {
    bc_copy(c, loc, 0, 0, rho);
    bc_copy(c, loc, 0, 0, u);
    bc_copy(c, loc, 0, 0, v);
    bc_copy(c, loc, 0, 0, p);
  } else {
    bc_copy(c, loc, 1, 0, rho);
    bc_copy(c, loc, 1, xdir, u);
    bc_copy(c, loc, 1, !xdir, v);
    bc_copy(c, loc, 1, 0, p);
  }
}

void or_apply_bc_single(const or_config *c, int type, int loc, double *rho, double *u, double *v, double *p) {
  /* boundary_condition_t%apply of one object; loc 0:+x 1:-x 2:+y 3:-y */
  bc_apply_one(c, type, loc, rho, u, v, p);
}

void or_apply_bcs(const or_config *c, double *rho, double *u, double *v, double *p) {
  /* flux_solver.f90:157-183: highest priority first; same priority in order +x, +y, -x, -y */
  int px = bc_priority(c->bc_plus_x), py = bc_priority(c->bc_plus_y);
  int mxp = bc_priority(c->bc_minus_x), myp = bc_priority(c->bc_minus_y);
  int maxp = px;
  if (py > maxp) maxp = py;
  if (mxp > maxp) maxp = mxp;
  if (myp > maxp) maxp = myp;
  for (int pr = maxp; pr >= 0; --pr) {
    if (px == pr) bc_apply_one(c, c->bc_plus_x, 0, rho, u, v, p);
    if (py == pr) bc_apply_one(c, c->bc_plus_y, 2, rho, u, v, p);
    if (mxp == pr) bc_apply_one(c, c->bc_minus_x, 1, rho, u, v, p);
    if (myp == pr) bc_apply_one(c, c->bc_minus_y, 3, rho, u, v, p);
  }
}

/* ------------------------------------------------------------------------- */
/* Reconstruction                                                              */
/* ------------------------------------------------------------------------- */
static inline double tvd_phi(int lim, double r) {
  switch (lim) {
    case OR_LIM_MINMOD: return fmax2(0.0, fmin2(r, 1.0));                        /* muscl_tvd2.fypp:196 */
    case OR_LIM_SUPERBEE: return fmax3(0.0, fmin2(2.0 * r, 1.0), fmin2(r, 2.0)); /* :198 */
    default: return (r + fabs(r)) / (1.0 + fabs(r));                             /* :200 */
  }
}

static void recon_tvd2(const or_config *c, int dir, const double *q, double *ev) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  OR_PAR
  for (int j = 0; j <= nj; ++j)
    for (int i = 0; i <= ni; ++i) {
      double q0, qm, q1, q2;
      if (dir == 0) { qm = q[CI(i - 1, j)]; q0 = q[CI(i, j)]; q1 = q[CI(i + 1, j)]; q2 = q[CI(i + 2, j)]; }
      else          { qm = q[CI(i, j - 1)]; q0 = q[CI(i, j)]; q1 = q[CI(i, j + 1)]; q2 = q[CI(i, j + 2)]; }
      double dm = q0 - qm, dp = q1 - q0, d3 = q2 - q1;
      if (fabs(dm) < EPS) dm = 0.0;
      if (fabs(dp) < EPS) dp = 0.0;
      if (fabs(d3) < EPS) d3 = 0.0;
      double r_L = dp / (dm + 1e-16);
      double r_R = dp / (d3 + 1e-16);
      double phi = tvd_phi(c->limiter, r_L);
      ev[EI(0, i, j)] = q0 + 0.5 * phi * dm;
      phi = tvd_phi(c->limiter, r_R);
      ev[EI(1, i, j)] = q1 - 0.5 * phi * d3;
    }
}

static void recon_mlp(const or_config *c, int dir, int order, const double *q, double *ev) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const double SMALL = 1e-30;
#define Q(a, b) q[CI((a), (b))]
  OR_PAR
  for (int j = 0; j <= nj; ++j)
    for (int i = 0; i <= ni; ++i) {
      double dm, dp, d3, dm3 = 0.0, d5 = 0.0;
      if (dir == 0) {
        dm = Q(i, j) - Q(i - 1, j); dp = Q(i + 1, j) - Q(i, j); d3 = Q(i + 2, j) - Q(i + 1, j);
      } else {
        dm = Q(i, j) - Q(i, j - 1); dp = Q(i, j + 1) - Q(i, j); d3 = Q(i, j + 2) - Q(i, j + 1);
      }
      if (fabs(dm) < EPS) dm = 0.0;
      if (fabs(dp) < EPS) dp = 0.0;
      if (fabs(d3) < EPS) d3 = 0.0;
      if (order == 5) {
        if (dir == 0) { dm3 = Q(i - 1, j) - Q(i - 2, j); d5 = Q(i + 3, j) - Q(i + 2, j); }
        else          { dm3 = Q(i, j - 1) - Q(i, j - 2); d5 = Q(i, j + 3) - Q(i, j + 2); }
        if (fabs(dm3) < EPS) dm3 = 0.0;
        if (fabs(d5) < EPS) d5 = 0.0;
      }
      double r_R_p1 = dp / (d3 + SMALL);
      double r_L = dp / (dm + SMALL);
      double beta_L, beta_R;
      if (order == 5) {
        double r_L_m1 = dm / (dm3 + SMALL);
        double r_L_p1 = d3 / (dp + SMALL);
        double r_R_p2 = d3 / (d5 + SMALL);
        double r_R = dm / (dp + SMALL);
        /* muscl_mlp.fypp:236-237 */
        beta_L = ((-(2.0 / (r_L_m1 + SMALL)) + 11.0) + 24.0 * r_L - 3.0 * r_L * r_L_p1) / 30.0;
        beta_R = ((-(2.0 / (r_R_p2 + SMALL)) + 11.0) + 24.0 * r_R_p1 - 3.0 * r_R_p1 * r_R) / 30.0;
      } else {
        beta_L = (1.0 + 2.0 * r_L) / 3.0;
        beta_R = (1.0 + 2.0 * r_R_p1) / 3.0;
      }
      double tt, tt1;
      if (dir == 0) {
        tt = fabs(Q(i, j + 1) - Q(i, j - 1)) / (fabs(Q(i + 1, j) - Q(i - 1, j)) + SMALL);
        tt1 = fabs(Q(i + 1, j + 1) - Q(i + 1, j - 1)) / (fabs(Q(i + 2, j) - Q(i, j)) + SMALL);
      } else {
        tt = fabs(Q(i + 1, j) - Q(i - 1, j)) / (fabs(Q(i, j + 1) - Q(i, j - 1)) + SMALL);
        tt1 = fabs(Q(i + 1, j + 1) - Q(i - 1, j + 1)) / (fabs(Q(i, j + 2) - Q(i, j)) + SMALL);
      }
      double aLt = (2.0 * fmax2(1.0, r_L) * (1.0 + fmax2(0.0, (tt1 / (r_R_p1 + SMALL))))) / (1.0 + tt);
      double aRt = (2.0 * fmax2(1.0, r_R_p1) * (1.0 + fmax2(0.0, (tt / (r_L + SMALL))))) / (1.0 + tt1);
      double aL = fmax2(1.0, fmin2(2.0, aLt));
      double aR = fmax2(1.0, fmin2(2.0, aRt));
      double q1 = (dir == 0) ? Q(i + 1, j) : Q(i, j + 1);
      ev[EI(0, i, j)] = Q(i, j) + 0.5 * fmax2(0.0, fmin3(aL * r_L, aL, beta_L)) * dm;
      ev[EI(1, i, j)] = q1 - 0.5 * fmax2(0.0, fmin3(aR * r_R_p1, aR, beta_R)) * d3;
    }
#undef Q
}

void or_interpolate_edges(const or_config *c, const double *q, double *i_edges, double *j_edges) {
  if (c->limiter <= OR_LIM_VAN_LEER) {
    recon_tvd2(c, 0, q, i_edges);
    recon_tvd2(c, 1, q, j_edges);
  } else {
    int order = c->limiter == OR_LIM_MLP3 ? 3 : 5;
    recon_mlp(c, 0, order, q, i_edges);
    recon_mlp(c, 1, order, q, j_edges);
  }
}

/* ------------------------------------------------------------------------- */
/* Edge fluxes                                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
  double rL, uL, vL, pL, rR, uR, vR, pR;
} edge_state;

/* SLAU family: slau_solver.fypp:332-493 */
static void flux_slau(int variant, double gamma, const edge_state *s, double n_x, double n_y, const double *pn6,
                      double out[4]) {
  const double THRESHOLD = 5e-3, C_SD1 = 0.01;
  double rho_L = s->rL, u_L = s->uL, v_L = s->vL, p_L = s->pL;
  double rho_R = s->rR, u_R = s->uR, v_R = s->vR, p_R = s->pR;
  double V_n_plus = u_L * n_x + v_L * n_y;
  double V_n_minus = u_R * n_x + v_R * n_y;
  double c_plus = sqrt(gamma * p_L / rho_L);
  double c_minus = sqrt(gamma * p_R / rho_R);
  double c_bar = 0.5 * (c_plus + c_minus);
  double V_n_bar = (rho_L * fabs(V_n_plus) + rho_R * fabs(V_n_minus)) / (rho_L + rho_R);
  double vel2 = sq(u_L) + sq(v_L) + sq(u_R) + sq(v_R);
  double M_hat = fmin2(1.0, sqrt(0.5 * vel2) / c_bar);
  double Chi = sq(1.0 - M_hat);
  double M_bar = V_n_bar / c_bar;
  double p_bar = 0.5 * (p_L + p_R);
  double rho_bar = 0.5 * (rho_L + rho_R);
  double M_plus = V_n_plus / c_bar;
  double M_minus = V_n_minus / c_bar;
  double Beta_plus, Beta_minus;
  if (fabs(M_plus) < 1.0) Beta_plus = 0.25 * (2.0 - M_plus) * sq(M_plus + 1.0);
  else Beta_plus = 0.5 * (1.0 + copysign(1.0, M_plus));
  if (fabs(M_minus) < 1.0) Beta_minus = 0.25 * (2.0 + M_minus) * sq(M_minus - 1.0);
  else Beta_minus = 0.5 * (1.0 + copysign(1.0, -M_minus));
  double delta_p = p_R - p_L;
  if (fabs(delta_p) < EPS || fabs(delta_p) < THRESHOLD * p_bar) delta_p = 0.0;
  double delta_rho = rho_R - rho_L;
  if (fabs(delta_rho) < EPS) delta_rho = 0.0;
  double delta_Beta = Beta_plus - Beta_minus;
  if (fabs(delta_Beta) < EPS) delta_Beta = 0.0;
  double p_L_minus_R = p_L - p_R;
  if (fabs(p_L_minus_R) < EPS) p_L_minus_R = 0.0;
  double theta = 0.0;
  int sd = (variant == OR_FLUX_SD_SLAU || variant == OR_FLUX_SD_SLAU2);
  if (sd) {
    double dps[6];
    for (int k = 0; k < 6; ++k) {
      dps[k] = pn6[k];
      if (dps[k] < EPS || dps[k] < THRESHOLD * p_bar) dps[k] = 0.0;
    }
    double mx6 = dps[0];
    for (int k = 1; k < 6; ++k) mx6 = fmax2(mx6, dps[k]);
    double delta_p_max = fmax2(fabs(delta_p), mx6);
    theta = fmin2(1.0, sq(((C_SD1 * fabs(delta_p) / p_bar) + C_SD1) / ((delta_p_max / p_bar) + C_SD1))) *
            fmax2(0.0, 1.0 - fabs(M_bar));
  }
  double p_tilde;
  if (variant == OR_FLUX_SLAU || variant == OR_FLUX_SD_SLAU)
    p_tilde = (p_bar + 0.5 * delta_Beta * p_L_minus_R) + (1.0 - Chi) * (Beta_plus + Beta_minus - 1.0) * p_bar;
  else
    p_tilde = p_bar + (0.5 * delta_Beta * p_L_minus_R) +
              sqrt(0.5 * vel2) * (Beta_plus + Beta_minus - 1.0) * rho_bar * c_bar;
  double g = -(fmax2(fmin2(M_plus, 0.0), -1.0) * fmin2(fmax2(M_minus, 0.0), 1.0));
  double mass_flux;
  double base = 0.5 * ((rho_L * V_n_plus) + (rho_R * V_n_minus) - fabs(V_n_bar) * delta_rho) * (1.0 - g);
  if (sd) mass_flux = base - ((theta / (2.0 * c_bar)) * delta_p);
  else mass_flux = base - ((Chi / (2.0 * c_bar)) * delta_p);
  double mfL = 0.5 * (mass_flux + fabs(mass_flux));
  double mfR = 0.5 * (mass_flux - fabs(mass_flux));
  double gg = gamma / (gamma - 1.0);
  double H_L = gg * (p_L / rho_L) + 0.5 * (sq(u_L) + sq(v_L));
  double H_R = gg * (p_R / rho_R) + 0.5 * (sq(u_R) + sq(v_R));
  out[0] = mfL + mfR;
  out[1] = (mfL * u_L) + (mfR * u_R) + p_tilde * n_x;
  out[2] = (mfL * v_L) + (mfR * v_R) + p_tilde * n_y;
  out[3] = (mfL * H_L) + (mfR * H_R);
}

/* AUSMPW+: ausmpw_plus_solver.fypp:335-460, 486-568 (alpha = 1/8) */
static double pw_P_plus(double M) {
  const double alpha = 1.0 / 8.0;
  if (fabs(M) > 1.0) return 0.5 * (1.0 + copysign(1.0, M));
  return 0.25 * sq(M + 1.0) * (2.0 - M) + alpha * M * sq(sq(M) - 1.0);
}
static double pw_P_minus(double M) {
  const double alpha = 1.0 / 8.0;
  if (fabs(M) > 1.0) return 0.5 * (1.0 - copysign(1.0, M));
  return 0.25 * sq(M - 1.0) * (2.0 + M) - alpha * M * sq(sq(M) - 1.0);
}
static double pw_M_plus(double M) {
  if (fabs(M) > 1.0) return 0.5 * (M + fabs(M));
  return 0.25 * sq(M + 1.0);
}
static double pw_M_minus(double M) {
  if (fabs(M) > 1.0) return 0.5 * (M - fabs(M));
  return -(0.25 * sq(M - 1.0));
}

static void flux_ausmpw(double gamma, int low_mach_fix, int dir, const edge_state *s, double n_x, double n_y,
                        const double *p4 /* p1_L,p1_R,p2_L,p2_R */, double out[4]) {
  const double LOW_MACH = 0.1;
  double rho_L = s->rL, u_LHS = s->uL, v_LHS = s->vL, p_L = s->pL;
  double rho_R = s->rR, u_RHS = s->uR, v_RHS = s->vR, p_R = s->pR;
  double U_L, U_R, V_L, V_R;
  if (dir == 0) { U_L = u_LHS; U_R = u_RHS; V_L = v_LHS; V_R = v_RHS; }
  else          { U_L = v_LHS; U_R = v_RHS; V_L = u_LHS; V_R = u_RHS; }
  double cs_L = sqrt(gamma * p_L / rho_L);
  double cs_R = sqrt(gamma * p_R / rho_R);
  double M_LHS = sqrt(sq(u_LHS) + sq(v_LHS)) / cs_L;
  double M_RHS = sqrt(sq(u_RHS) + sq(v_RHS)) / cs_R;
  if (low_mach_fix && (fabs(M_LHS) < LOW_MACH || fabs(M_RHS) < LOW_MACH)) {
    double f_M = fmin2(1.0, fmax2(M_LHS, M_RHS));
    double ulf = 0.5 * (u_LHS + u_RHS) + f_M * 0.5 * (u_LHS - u_RHS);
    double urf = 0.5 * (u_LHS + u_RHS) - f_M * 0.5 * (u_LHS - u_RHS);
    double vlf = 0.5 * (v_LHS + v_RHS) + f_M * 0.5 * (v_LHS - v_RHS);
    double vrf = 0.5 * (v_LHS + v_RHS) - f_M * 0.5 * (v_LHS - v_RHS);
    u_LHS = ulf; u_RHS = urf; v_LHS = vlf; v_RHS = vrf;
  }
  double gg = gamma / (gamma - 1.0);
  double H_L = gg * (p_L / rho_L) + 0.5 * (sq(u_LHS) + sq(v_LHS));
  double H_R = gg * (p_R / rho_R) + 0.5 * (sq(u_RHS) + sq(v_RHS));
  double H_normal = fmin2(H_L - 0.5 * sq(V_L), H_R - 0.5 * sq(V_R));
  double c_s = sqrt(2.0 * ((gamma - 1.0) / (gamma + 1.0)) * H_normal);
  double c_half;
  if (0.5 * (U_L + U_R) < 0.0) c_half = sq(c_s) / fmax2(fabs(U_R), c_s);
  else c_half = sq(c_s) / fmax2(fabs(U_L), c_s);
  double M_L = U_L / c_half;
  double M_R = U_R / c_half;
  double M_L_plus = pw_M_plus(M_L);
  double M_R_minus = pw_M_minus(M_R);
  double P_R_minus = pw_P_minus(M_R);
  double P_L_plus = pw_P_plus(M_L);
  double p_s = p_L * P_L_plus + p_R * P_R_minus;
  /* get_w (ausmpw_plus_solver.fypp:558-568) */
  double mt = fmin2((p_L / p_R), (p_R / p_L));
  double w = 1.0 - mt * mt * mt;
  /* get_f (ausmpw_plus_solver.fypp:532-556) */
  double f_L, f_R;
  if (fabs(p_s) < TINY) {
    f_L = 0.0;
    f_R = 0.0;
  } else {
    double mn4 = p4[0];
    for (int k = 1; k < 4; ++k) mn4 = fmin2(mn4, p4[k]);
    double m1 = fmin2(1.0, mn4 / fmin2(p_L, p_R));
    double min_term = sq(m1);
    f_L = ((p_L / p_s) - 1.0) * min_term;
    f_R = ((p_R / p_s) - 1.0) * min_term;
  }
  double Mb_L, Mb_R;
  if (M_L_plus + M_R_minus < 0.0) {
    Mb_L = M_L_plus * w * (1.0 + f_L);
    Mb_R = M_R_minus + M_L_plus * ((1.0 - w) * (1.0 + f_L) - f_R);
  } else {
    Mb_L = M_L_plus + M_R_minus * ((1.0 - w) * (1.0 + f_R) - f_L);
    Mb_R = M_R_minus * w * (1.0 + f_R);
  }
  double mfL = Mb_L * c_half * rho_L;
  double mfR = Mb_R * c_half * rho_R;
  out[0] = mfL + mfR;
  out[1] = (mfL * u_LHS) + (mfR * u_RHS) + ((P_L_plus * n_x * p_L) + (P_R_minus * n_x * p_R));
  out[2] = (mfL * v_LHS) + (mfR * v_RHS) + ((P_L_plus * n_y * p_L) + (P_R_minus * n_y * p_R));
  out[3] = (mfL * H_L) + (mfR * H_R);
}

/* M-AUSMPW+: m_ausmpw_plus_solver.fypp:437-738 (the code path the reference runs: AUSMPW+
 * pressure splitting without the alpha term, w from the normal pressure ratio), get_f 829-862 with
 * its 1e-10 flushes, phi_L_half / phi_R_half 790-827 blending the limiter-of-choice states with the
 * superbee states, and the 1e3*epsilon flushes of the assembled flux 715-735. */
static double m_phi_L_half(double phi_L, double phi_R, double phi_L_sb, double a) {
  if (fabs(phi_R - phi_L) < 1.0 || fabs(phi_L_sb - phi_L) < EPS) return phi_L;
  return phi_L + fmax2(0.0, (phi_R - phi_L) * (phi_L_sb - phi_L)) / ((phi_R - phi_L) * fabs(phi_L_sb - phi_L)) *
                     fmin3(a, 0.5 * fabs(phi_R - phi_L), fabs(phi_L_sb - phi_L));
}

static double m_phi_R_half(double phi_L, double phi_R, double phi_R_sb, double a) {
  if (fabs(phi_L - phi_R) < EPS || fabs(phi_R_sb - phi_R) < 1.0) return phi_R;
  return phi_R + fmax2(0.0, (phi_L - phi_R) * (phi_R_sb - phi_R)) / ((phi_L - phi_R) * fabs(phi_R_sb - phi_R)) *
                     fmin3(a, 0.5 * fabs(phi_L - phi_R), fabs(phi_R_sb - phi_R));
}

static void flux_m_ausmpw(double gamma, int low_mach_fix, int dir, const edge_state *s, const edge_state *sb,
                          double n_x, double n_y, const double *p4 /* p1_L,p1_R,p2_L,p2_R */, double out[4]) {
  const double LOW_MACH = 0.1;
  const double FEPS = EPS * 1e3;
  double rho_L = s->rL, u_LHS = s->uL, v_LHS = s->vL, p_L = s->pL;
  double rho_R = s->rR, u_RHS = s->uR, v_RHS = s->vR, p_R = s->pR;
  double U_L, U_R, V_L, V_R;
  if (dir == 0) { U_L = u_LHS; U_R = u_RHS; V_L = v_LHS; V_R = v_RHS; }
  else          { U_L = v_LHS; U_R = v_RHS; V_L = u_LHS; V_R = u_RHS; }
  double cs_L = sqrt(gamma * p_L / rho_L);
  double cs_R = sqrt(gamma * p_R / rho_R);
  double M_LHS = sqrt(sq(u_LHS) + sq(v_LHS)) / cs_L;
  double M_RHS = sqrt(sq(u_RHS) + sq(v_RHS)) / cs_R;
  if (low_mach_fix && (fabs(M_LHS) < LOW_MACH || fabs(M_RHS) < LOW_MACH)) {
    double f_M = fmin2(1.0, fmax2(M_LHS, M_RHS));
    double ulf = 0.5 * (u_LHS + u_RHS) + f_M * 0.5 * (u_LHS - u_RHS);
    double urf = 0.5 * (u_LHS + u_RHS) - f_M * 0.5 * (u_LHS - u_RHS);
    double vlf = 0.5 * (v_LHS + v_RHS) + f_M * 0.5 * (v_LHS - v_RHS);
    double vrf = 0.5 * (v_LHS + v_RHS) - f_M * 0.5 * (v_LHS - v_RHS);
    u_LHS = ulf; u_RHS = urf; v_LHS = vlf; v_RHS = vrf;
  }
  double H_L = (gamma / (gamma - 1.0)) * (p_L / rho_L) + 0.5 * (sq(u_LHS) + sq(v_LHS));
  double H_R = (gamma / (gamma - 1.0)) * (p_R / rho_R) + 0.5 * (sq(u_RHS) + sq(v_RHS));
  double H_normal = fmin2(H_L - 0.5 * sq(V_L), H_R - 0.5 * sq(V_R));
  double c_s = sqrt(2.0 * ((gamma - 1.0) / (gamma + 1.0)) * H_normal);
  double c_half;
  if (0.5 * (U_L + U_R) < 0.0) c_half = sq(c_s) / fmax2(fabs(U_R), c_s);
  else c_half = sq(c_s) / fmax2(fabs(U_L), c_s);
  double M_L = U_L / c_half;
  double M_R = U_R / c_half;
  double M_L_plus = (fabs(M_L) > 1.0) ? 0.5 * (M_L + fabs(M_L)) : 0.25 * sq(M_L + 1.0);
  double M_R_minus = (fabs(M_R) > 1.0) ? 0.5 * (M_R - fabs(M_R)) : -(0.25 * sq(M_R - 1.0));
  double P_R_minus = (fabs(M_R) > 1.0) ? 0.5 * (1.0 - copysign(1.0, M_R)) : 0.25 * sq(M_R - 1.0) * (2.0 + M_R);
  double P_L_plus = (fabs(M_L) > 1.0) ? 0.5 * (1.0 + copysign(1.0, M_L)) : 0.25 * sq(M_L + 1.0) * (2.0 - M_L);
  double p_s = p_L * P_L_plus + p_R * P_R_minus;
  double mt = fmin2((p_L / p_R), (p_R / p_L));
  double w = 1.0 - (mt * mt * mt);
  double f_L, f_R;
  if (fabs(p_s) < TINY) {
    f_L = 0.0;
    f_R = 0.0;
  } else {
    const double FE = 1e-10;
    double mn4 = fmin2(fmin2(fmin2(p4[0], p4[1]), p4[2]), p4[3]);
    double min_term = sq(fmin2(1.0, mn4 / fmin2(p_L, p_R)));
    if (fabs(1.0 - min_term) < FE) min_term = 1.0;
    double fl = (p_L / p_s) - 1.0;
    if (fabs(fl) < FE) fl = 0.0;
    f_L = fl * min_term;
    double fr = (p_R / p_s) - 1.0;
    if (fabs(fr) < FE) fr = 0.0;
    f_R = fr * min_term;
  }
  double Mb_L, Mb_R;
  if (M_L_plus + M_R_minus < 0.0) {
    Mb_L = M_L_plus * w * (1.0 + f_L);
    Mb_R = M_R_minus + M_L_plus * ((1.0 - w) * (1.0 + f_L) - f_R);
  } else {
    Mb_L = M_L_plus + M_R_minus * ((1.0 - w) * (1.0 + f_R) - f_L);
    Mb_R = M_R_minus * w * (1.0 + f_R);
  }
  double a = 1.0 - sq(fmin2(1.0, fmax2(fabs(M_L), fabs(M_R))));
  double rho_L_half = m_phi_L_half(rho_L, rho_R, sb->rL, a);
  double rho_R_half = m_phi_R_half(rho_L, rho_R, sb->rR, a);
  double u_L_half = m_phi_L_half(u_LHS, u_RHS, sb->uL, a);
  double u_R_half = m_phi_R_half(u_LHS, u_RHS, sb->uR, a);
  double v_L_half = m_phi_L_half(v_LHS, v_RHS, sb->vL, a);
  double v_R_half = m_phi_R_half(v_LHS, v_RHS, sb->vR, a);
  double p_L_half = m_phi_L_half(p_L, p_R, sb->pL, a);
  double p_R_half = m_phi_R_half(p_L, p_R, sb->pR, a);
  double H_L_half = (gamma / (gamma - 1.0)) * (p_L_half / rho_L_half) + 0.5 * (sq(u_L_half) + sq(v_L_half));
  double H_R_half = (gamma / (gamma - 1.0)) * (p_R_half / rho_R_half) + 0.5 * (sq(u_R_half) + sq(v_R_half));
  double mfL = Mb_L * c_half * rho_L_half;
  double mfR = Mb_R * c_half * rho_R_half;
  if (fabs(mfL + mfR) < FEPS) {
    mfL = 0.0;
    mfR = 0.0;
  }
  out[0] = mfL + mfR;
  double pu = (P_L_plus * n_x * p_L) + (P_R_minus * n_x * p_R);
  double mu = (mfL * u_L_half) + (mfR * u_R_half);
  out[1] = (fabs(mu) < FEPS) ? pu : mu + pu;
  double pv = (P_L_plus * n_y * p_L) + (P_R_minus * n_y * p_R);
  double mv = (mfL * v_L_half) + (mfR * v_R_half);
  out[2] = (fabs(mv) < FEPS) ? pv : mv + pv;
  out[3] = (mfL * H_L_half) + (mfR * H_R_half);
  if (fabs(out[3]) < FEPS) out[3] = 0.0;
}

/* Roe (reference variant): roe_solver.fypp:420-511, f_h 519-527 */
static double roe_f_h(double pL, double pR) {
  double ratio = fabs(pR - pL) / (fmin2(pR, pL) + 1e-16);
  return (ratio < 1.0) ? 0.0 : 1.0;
}

static void flux_roe(double gamma, const edge_state *s, double nx, double ny, const double *pf /* 5 pairs */,
                     double out[4]) {
  double rl = fabs(s->rL), ul = s->uL, vl = s->vL, pl = fabs(s->pL);
  double gg = gamma / (gamma - 1.0);
  double hl = gg * (pl / rl) + 0.5 * (sq(ul) + sq(vl));
  double rr = fabs(s->rR), ur = s->uR, vr = s->vR, pr = fabs(s->pR);
  double hr = gg * (pr / rr) + 0.5 * (sq(ur) + sq(vr));
  double rav = sqrt(rl * rr);
  double gam1 = gamma - 1.0;
  double dd = rav / rl;
  double dd1 = 1.0 / (1.0 + dd);
  double uav = (ul + dd * ur) * dd1;
  double vav = (vl + dd * vr) * dd1;
  double hav = (hl + dd * hr) * dd1;
  double q2a = 0.5 * (uav * uav + vav * vav);
  double c2a = gam1 * (hav - q2a);
  double cav = sqrt(c2a);
  double uv = uav * nx + vav * ny;
  double du = (ur - ul) * nx + (vr - vl) * ny;
  double du_trans = -((ur - ul) * ny) + (vr - vl) * nx;
  double cs_R = sqrt(gamma * pr / rr);
  double cs_L = sqrt(gamma * pl / rl);
  double eabs1 = fabs(fmin2(uv - cav, (ul * nx + vl * ny) - cs_L));
  double eabs2 = fabs(uv);
  double eabs4 = fabs(fmax2(uv + cav, (ur * nx + vr * ny) + cs_R));
  double h1 = rav * cav * du;
  double h2 = eabs1 * (pr - pl - h1) / (2.0 * c2a);
  double h3 = eabs2 * (rr - rl - (pr - pl) / c2a);
  double h4 = eabs2 * rav;
  double h5 = eabs4 * (pr - pl + h1) / (2.0 * c2a);
  double psi = fabs(cs_R - cs_L) / fmax2(cs_R, cs_L);
  double FR[4] = {rr * uv, rr * ur * uv + nx * pr, rr * vr * uv + ny * pr, rr * hr * uv};
  double FL[4] = {rl * uv, rl * ul * uv + nx * pl, rl * vl * uv + ny * pl, rl * hl * uv};
  double rm[4];
  rm[0] = (h2 + h3 + h5);
  rm[1] = (h2 * (uav - cav * nx) + h3 * uav + h4 * (ur - ul - du * nx) + h5 * (uav + cav * nx));
  rm[2] = (h2 * (vav - cav * ny) + h3 * vav + h4 * (vr - vl - du * ny) + h5 * (vav + cav * ny));
  rm[3] = (h2 * (hav - cav * uv) + h3 * q2a + h4 * (uav * (ur - ul) + vav * (vr - vl) - uv * du) +
           h5 * (hav + cav * uv));
  for (int k = 0; k < 4; ++k) out[k] = 0.5 * (FR[k] + FL[k] - rm[k]);
  double fp = 0.0;
  for (int k = 0; k < 5; ++k) {
    double f = roe_f_h(pf[2 * k], pf[2 * k + 1]);
    fp = (k == 0) ? f : fmax2(fp, f);
  }
  out[1] = out[1] - 0.5 * fp * rav * cav * (-(du_trans * ny) + psi * du * nx);
  out[2] = out[2] - 0.5 * fp * rav * cav * (du_trans * nx + psi * du * ny);
}

/* AUSM+ assembled from interface_state (ausm_plus_solver.f90:394-643) with the
 * 'AUSM+-u' flags (M_p = p_u = 0), beta = 1/8 (mod_input.f90:139), alpha = 3/16
 * (ausm_plus_solver.f90:51), and the upwind rule of the commented solve (312-325). */
static void flux_ausm_plus(double gamma, const edge_state *s, double n_x, double n_y, double out[4]) {
  const double beta = 1.0 / 8.0, alpha = 3.0 / 16.0;
  double rho_L = s->rL, u_L = s->uL, v_L = s->vL, p_L = s->pL;
  double rho_R = s->rR, u_R = s->uR, v_R = s->vR, p_R = s->pR;
  double gamma_factor = ((2.0 * (gamma - 1.0)) / (gamma + 1.0));
  double vel_L = u_L * n_x + v_L * n_y;
  double vel_R = u_R * n_x + v_R * n_y;
  double gg = gamma / (gamma - 1.0);
  double H_L = (p_L / rho_L) * gg + 0.5 * (u_L * u_L + v_L * v_L);
  double H_R = (p_R / rho_R) * gg + 0.5 * (u_R * u_R + v_R * v_R);
  double a_crit_L = sqrt(gamma_factor * fabs(H_L));
  double a_crit_R = sqrt(gamma_factor * fabs(H_R));
  double a_circ_L = sq(a_crit_L) / fmax2(a_crit_L, vel_L);
  double a_circ_R = sq(a_crit_R) / fmax2(a_crit_R, -vel_R);
  double a_half = fmin2(a_circ_L, a_circ_R);
  double M_L = vel_L / a_half;
  double M_R = vel_R / a_half;
  /* split_mach_deg_4 */
  double M2p_L = 0.25 * sq(M_L + 1.0), M2m_L = -(0.25 * sq(M_L - 1.0));
  double M2p_R = 0.25 * sq(M_R + 1.0), M2m_R = -(0.25 * sq(M_R - 1.0));
  double M_plus = (fabs(M_L) >= 1.0) ? 0.5 * (M_L + fabs(M_L)) : M2p_L * (1.0 - 16.0 * beta * M2m_L);
  double M_minus = (fabs(M_R) >= 1.0) ? 0.5 * (M_R - fabs(M_R)) : M2m_R * (1.0 + 16.0 * beta * M2p_R);
  double M_p = 0.0;
  double M_sum = M_plus + M_minus;
  if (fabs(M_sum) < 1e-13) M_sum = 0.0;
  double M_half = M_sum + M_p;
  if (fabs(M_half) < 1e-13) M_half = 0.0;
  /* split_pressure_deg_5 */
  double P_plus = (fabs(M_L) >= 1.0) ? (0.5 * (M_L + fabs(M_L))) / M_L
                                     : M2p_L * ((2.0 - M_L) - 16.0 * alpha * M_L * M2m_L);
  double P_minus = (fabs(M_R) >= 1.0) ? (0.5 * (M_R - fabs(M_R))) / M_R
                                      : M2m_R * ((-2.0 - M_R) + 16.0 * alpha * M_R * M2p_R);
  double p_u = 0.0;
  double p_half = P_plus * p_L + P_minus * p_R + p_u;
  double mass_flux;
  if (M_half > 0.0) {
    mass_flux = M_half * a_half * rho_L;
    out[0] = mass_flux;
    out[1] = mass_flux * u_L + (n_x * p_half);
    out[2] = mass_flux * v_L + (n_y * p_half);
    out[3] = mass_flux * H_L;
  } else {
    mass_flux = M_half * a_half * rho_R;
    out[0] = mass_flux;
    out[1] = mass_flux * u_R + (n_x * p_half);
    out[2] = mass_flux * v_R + (n_y * p_half);
    out[3] = mass_flux * H_R;
  }
}

/* HLLC (Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
 * 2009, sec. 10.4-10.6): PVRS pressure estimate, pressure-based wave speeds
 * (eq. 10.59-10.60), contact speed eq. 10.37, star states eq. 10.39.  The
 * reference hllc_solver.fypp:345-353 sketches exactly this (p_pvrs, p_star,
 * S_L = u_L - cs_L*q_L, ...) but is truncated, so parity is pinned by what Toro's scheme fixes:
 * supersonic upwinding, exact isolated contacts and shear waves, and Toro's alternative
 * star-flux form (eq. 10.41-10.42) on general states. */
static void flux_hllc(double gamma, const edge_state *s, double n_x, double n_y, double out[4]) {
  double rL = s->rL, uL = s->uL, vL = s->vL, pL = s->pL;
  double rR = s->rR, uR = s->uR, vR = s->vR, pR = s->pR;
  double gm1 = gamma - 1.0;
  double qL = uL * n_x + vL * n_y;
  double qR = uR * n_x + vR * n_y;
  double cL = sqrt(gamma * pL / rL);
  double cR = sqrt(gamma * pR / rR);
  double rho_bar = 0.5 * (rL + rR);
  double cs_bar = 0.5 * (cL + cR);
  double p_pvrs = 0.5 * (pL + pR) - 0.5 * (qR - qL) * rho_bar * cs_bar;
  double p_star = fmax2(0.0, p_pvrs);
  double gfac = (gamma + 1.0) / (2.0 * gamma);
  double wL = (p_star <= pL) ? 1.0 : sqrt(1.0 + gfac * (p_star / pL - 1.0));
  double wR = (p_star <= pR) ? 1.0 : sqrt(1.0 + gfac * (p_star / pR - 1.0));
  double SL = qL - cL * wL;
  double SR = qR + cR * wR;
  double mL = rL * (SL - qL);
  double mR = rR * (SR - qR);
  double S_star = (pR - pL + qL * mL - qR * mR) / (mL - mR);
  double EL = pL / gm1 + 0.5 * rL * (uL * uL + vL * vL);
  double ER = pR / gm1 + 0.5 * rR * (uR * uR + vR * vR);
  if (0.0 <= SL) {
    out[0] = rL * qL;
    out[1] = rL * uL * qL + pL * n_x;
    out[2] = rL * vL * qL + pL * n_y;
    out[3] = qL * (EL + pL);
  } else if (SR <= 0.0) {
    out[0] = rR * qR;
    out[1] = rR * uR * qR + pR * n_x;
    out[2] = rR * vR * qR + pR * n_y;
    out[3] = qR * (ER + pR);
  } else if (S_star >= 0.0) {
    double fac = mL / (SL - S_star);
    double Us0 = fac;
    double Us1 = fac * (uL + (S_star - qL) * n_x);
    double Us2 = fac * (vL + (S_star - qL) * n_y);
    double Us3 = fac * (EL / rL + (S_star - qL) * (S_star + pL / mL));
    out[0] = rL * qL + SL * (Us0 - rL);
    out[1] = (rL * uL * qL + pL * n_x) + SL * (Us1 - rL * uL);
    out[2] = (rL * vL * qL + pL * n_y) + SL * (Us2 - rL * vL);
    out[3] = qL * (EL + pL) + SL * (Us3 - EL);
  } else {
    double fac = mR / (SR - S_star);
    double Us0 = fac;
    double Us1 = fac * (uR + (S_star - qR) * n_x);
    double Us2 = fac * (vR + (S_star - qR) * n_y);
    double Us3 = fac * (ER / rR + (S_star - qR) * (S_star + pR / mR));
    out[0] = rR * qR + SR * (Us0 - rR);
    out[1] = (rR * uR * qR + pR * n_x) + SR * (Us1 - rR * uR);
    out[2] = (rR * vR * qR + pR * n_y) + SR * (Us2 - rR * vR);
    out[3] = qR * (ER + pR) + SR * (Us3 - ER);
  }
}

void or_edge_fluxes(const or_config *c, const or_geometry *g, const double *rho, const double *u, const double *v,
                    const double *p, const double *edges, double *iflux, double *jflux) {
  const size_t ne = (size_t)2 * (c->ni + 1) * (c->nj + 1);
  or_edge_fluxes_sb(c, g, rho, u, v, p, edges, c->flux == OR_FLUX_M_AUSMPW_PLUS ? edges + 8 * ne : NULL, iflux,
                    jflux);
}

void or_edge_fluxes_sb(const or_config *c, const or_geometry *g, const double *rho, const double *u, const double *v,
                       const double *p, const double *edges, const double *edges_sb, double *iflux, double *jflux) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const size_t ne = (size_t)2 * (ni + 1) * (nj + 1);
  const double *rho_i = edges + 0 * ne, *rho_j = edges + 1 * ne;
  const double *u_i = edges + 2 * ne, *u_j = edges + 3 * ne;
  const double *v_i = edges + 4 * ne, *v_j = edges + 5 * ne;
  const double *p_i = edges + 6 * ne, *p_j = edges + 7 * ne;
  (void)rho;
  (void)u;
  (void)v;
#define P(a, b) p[CI((a), (b))]
  for (int dir = 0; dir < 2; ++dir) {
    const double *er = dir == 0 ? rho_i : rho_j, *eu = dir == 0 ? u_i : u_j;
    const double *evv = dir == 0 ? v_i : v_j, *ep = dir == 0 ? p_i : p_j;
    int i0 = dir == 0 ? 0 : 1, j0 = dir == 0 ? 1 : 0;
    OR_PAR
    for (int j = j0; j <= nj; ++j)
      for (int i = i0; i <= ni; ++i) {
        edge_state s = {er[EI(0, i, j)], eu[EI(0, i, j)], evv[EI(0, i, j)], ep[EI(0, i, j)],
                        er[EI(1, i, j)], eu[EI(1, i, j)], evv[EI(1, i, j)], ep[EI(1, i, j)]};
        edge_state sb = s;
        if (edges_sb) {
          const double *b = edges_sb;
          const size_t o0 = (size_t)(dir == 0 ? 0 : 1) * ne;  /* rho_i | rho_j, then u, v, p pairs */
          sb.rL = b[o0 + 0 * 2 * ne + EI(0, i, j)]; sb.rR = b[o0 + 0 * 2 * ne + EI(1, i, j)];
          sb.uL = b[o0 + 1 * 2 * ne + EI(0, i, j)]; sb.uR = b[o0 + 1 * 2 * ne + EI(1, i, j)];
          sb.vL = b[o0 + 2 * 2 * ne + EI(0, i, j)]; sb.vR = b[o0 + 2 * 2 * ne + EI(1, i, j)];
          sb.pL = b[o0 + 3 * 2 * ne + EI(0, i, j)]; sb.pR = b[o0 + 3 * 2 * ne + EI(1, i, j)];
        }
        int e = dir == 0 ? 1 : 2; /* RIGHT_IDX / TOP_IDX */
        size_t cc = CI(i, j);
        double nx = g->edge_norm[8 * cc + 2 * e + 0];
        double ny = g->edge_norm[8 * cc + 2 * e + 1];
        double out[4];
        switch (c->flux) {
          case OR_FLUX_SLAU:
          case OR_FLUX_SLAU2:
          case OR_FLUX_SD_SLAU:
          case OR_FLUX_SD_SLAU2: {
            double d6[6];
            if (dir == 0) {
              d6[0] = fabs(P(i, j + 1) - P(i, j));
              d6[1] = fabs(P(i - 1, j) - P(i, j));
              d6[2] = fabs(P(i, j - 1) - P(i, j));
              d6[3] = fabs(P(i + 1, j - 1) - P(i + 1, j));
              d6[4] = fabs(P(i + 2, j) - P(i + 1, j));
              d6[5] = fabs(P(i + 1, j + 1) - P(i + 1, j));
            } else {
              d6[0] = fabs(P(i - 1, j) - P(i, j));
              d6[1] = fabs(P(i, j - 1) - P(i, j));
              d6[2] = fabs(P(i + 1, j) - P(i, j));
              d6[3] = fabs(P(i + 1, j + 1) - P(i, j + 1));
              d6[4] = fabs(P(i, j + 2) - P(i, j + 1));
              d6[5] = fabs(P(i - 1, j + 1) - P(i, j + 1));
            }
            flux_slau(c->flux, c->gamma, &s, nx, ny, d6, out);
            break;
          }
          case OR_FLUX_AUSMPW_PLUS: {
            double p4[4];
            if (dir == 0) { p4[0] = P(i, j - 1); p4[1] = P(i + 1, j - 1); p4[2] = P(i, j + 1); p4[3] = P(i + 1, j + 1); }
            else          { p4[0] = P(i + 1, j); p4[1] = P(i + 1, j + 1); p4[2] = P(i - 1, j); p4[3] = P(i - 1, j + 1); }
            flux_ausmpw(c->gamma, c->apply_low_mach_fix, dir, &s, nx, ny, p4, out);
            break;
          }
          case OR_FLUX_ROE: {
            double pf[10];
            if (dir == 0) {
              pf[0] = P(i, j); pf[1] = P(i + 1, j);
              pf[2] = P(i, j); pf[3] = P(i, j + 1);
              pf[4] = P(i, j); pf[5] = P(i, j - 1);
              pf[6] = P(i, j + 1); pf[7] = P(i + 1, j + 1);
              pf[8] = P(i, j - 1); pf[9] = P(i + 1, j - 1);
            } else {
              pf[0] = P(i, j); pf[1] = P(i, j + 1);
              pf[2] = P(i, j); pf[3] = P(i - 1, j);
              pf[4] = P(i, j); pf[5] = P(i + 1, j);
              pf[6] = P(i, j + 1); pf[7] = P(i - 1, j + 1);
              pf[8] = P(i, j + 1); pf[9] = P(i + 1, j + 1);
            }
            flux_roe(c->gamma, &s, nx, ny, pf, out);
            break;
          }
          case OR_FLUX_M_AUSMPW_PLUS: {
            double p4[4];
            if (dir == 0) { p4[0] = P(i, j - 1); p4[1] = P(i + 1, j - 1); p4[2] = P(i, j + 1); p4[3] = P(i + 1, j + 1); }
            else          { p4[0] = P(i + 1, j); p4[1] = P(i + 1, j + 1); p4[2] = P(i - 1, j); p4[3] = P(i - 1, j + 1); }
            flux_m_ausmpw(c->gamma, c->apply_low_mach_fix, dir, &s, &sb, nx, ny, p4, out);
            break;
          }
          case OR_FLUX_AUSM_PLUS: flux_ausm_plus(c->gamma, &s, nx, ny, out); break;
          default: flux_hllc(c->gamma, &s, nx, ny, out); break;
        }
        double *dst = dir == 0 ? &iflux[FI(0, i, j)] : &jflux[FJ(0, i, j)];
        for (int k = 0; k < 4; ++k) dst[k] = out[k];
      }
  }
#undef P
}

/* ------------------------------------------------------------------------- */
/* Divergence: flux_solver.f90:187-459                                        */
/* ------------------------------------------------------------------------- */
void or_flux_split_edges(const or_config *c, const or_geometry *g, const double *iflux, const double *jflux,
                         double *d_rho, double *d_rho_u, double *d_rho_v, double *d_rho_E) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  const double REL_TOL = 1e-14;
  const double NEAR_ZERO = 5.0 * TINY;
  double *dst[4] = {d_rho, d_rho_u, d_rho_v, d_rho_E};
  const size_t ncell = (size_t)(ni + 2 * nh) * (nj + 2 * nh);
  for (int v = 0; v < 4; ++v) memset(dst[v], 0, ncell * sizeof(double));
  for (int v = 0; v < 4; ++v)
    OR_PAR
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t cc = CI(i, j);
        const double *dl = &g->edge_lengths[4 * cc];
        double Fi = iflux[FI(v, i, j)], Fim = iflux[FI(v, i - 1, j)];
        double Gj = jflux[FJ(v, i, j)], Gjm = jflux[FJ(v, i, j - 1)];
        double mi = fmax3(fabs(Fi), fabs(Fim), 1.0);
        double mj = fmax3(fabs(Gj), fabs(Gjm), 1.0);
        double ri = ((Fi / mi * dl[1]) + (-(Fim / mi * dl[3]))) * mi;
        double rj = ((Gj / mj * dl[2]) + (-(Gjm / mj * dl[0]))) * mj;
        if (fabs(ri) < mi * REL_TOL) ri = 0.0;
        if (fabs(rj) < mj * REL_TOL) rj = 0.0;
        double fl = ri + rj;
        if (fabs(fl) < NEAR_ZERO) fl = 0.0;
        dst[v][cc] = -fl;
      }
  for (int v = 0; v < 4; ++v)
    OR_PAR
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t cc = CI(i, j);
        dst[v][cc] = dst[v][cc] / g->volume[cc];
      }
}

void or_solve(const or_config *c, const or_geometry *g, double *rho, double *u, double *v, double *p,
              double *d_rho, double *d_rho_u, double *d_rho_v, double *d_rho_E) {
  const int ni = c->ni, nj = c->nj;
  or_apply_bcs(c, rho, u, v, p);
  const size_t ne = (size_t)2 * (ni + 1) * (nj + 1);
  double *edges = (double *)calloc(8 * ne, sizeof(double));
  or_interpolate_edges(c, rho, edges + 0 * ne, edges + 1 * ne);
  or_interpolate_edges(c, u, edges + 2 * ne, edges + 3 * ne);
  or_interpolate_edges(c, v, edges + 4 * ne, edges + 5 * ne);
  or_interpolate_edges(c, p, edges + 6 * ne, edges + 7 * ne);
  double *iflux = (double *)calloc((size_t)4 * (ni + 1) * nj, sizeof(double));
  double *jflux = (double *)calloc((size_t)4 * ni * (nj + 1), sizeof(double));
  double *edges_sb = NULL;
  if (c->flux == OR_FLUX_M_AUSMPW_PLUS) {
    /* second reconstruction with the superbee limiter (m_ausmpw_plus_solver.fypp:206-211) */
    or_config csb = *c;
    csb.limiter = OR_LIM_SUPERBEE;
    edges_sb = (double *)calloc(8 * ne, sizeof(double));
    or_interpolate_edges(&csb, rho, edges_sb + 0 * ne, edges_sb + 1 * ne);
    or_interpolate_edges(&csb, u, edges_sb + 2 * ne, edges_sb + 3 * ne);
    or_interpolate_edges(&csb, v, edges_sb + 4 * ne, edges_sb + 5 * ne);
    or_interpolate_edges(&csb, p, edges_sb + 6 * ne, edges_sb + 7 * ne);
  }
  or_edge_fluxes_sb(c, g, rho, u, v, p, edges, edges_sb, iflux, jflux);
  free(edges_sb);
  or_flux_split_edges(c, g, iflux, jflux, d_rho, d_rho_u, d_rho_v, d_rho_E);
  free(edges);
  free(iflux);
  free(jflux);
}

/* ------------------------------------------------------------------------- */
/* Time integration: mod_fluid.f90                                            */
/* ------------------------------------------------------------------------- */
static size_t ncell(const or_config *c) { return (size_t)(c->ni + 2 * c->nh) * (c->nj + 2 * c->nh); }

void or_fluid_init_from_prims(const or_config *c, or_fluid *f) {
  or_prim_to_cons(c, f->rho, f->u, f->v, f->p, f->rho_u, f->rho_v, f->rho_E);
  or_sound_speed(c, f->rho, f->p, f->cs);
}

double or_get_timestep(const or_config *c, const or_geometry *g, const or_fluid *f) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  double m = 0.0;
  int first = 1;
#ifdef ORACLE_OMP
  /* timing build: the same minimum as the serial scan whenever no value is NaN */
  m = INFINITY;
  first = 0;
#pragma omp parallel for schedule(static) reduction(min : m)
#endif
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i) {
      size_t k = CI(i, j);
      double val = c->cfl / (((fabs(f->u[k]) + f->cs[k]) / g->dx[k]) + ((fabs(f->v[k]) + f->cs[k]) / g->dy[k]));
      if (first || val < m) m = val;
      first = 0;
    }
  return m;
}

/* per-stage state: conserved + primitive fields of one fluid_t */
typedef struct {
  double *a[8];
} stage_buf;

static void alloc_stage(const or_config *c, stage_buf *s) {
  for (int k = 0; k < 8; ++k) s->a[k] = (double *)calloc(ncell(c), sizeof(double));
}
static void free_stage(stage_buf *s) {
  for (int k = 0; k < 8; ++k) free(s->a[k]);
}
static void fluid_to_stage(const or_fluid *f, stage_buf *s, size_t n) {
  const double *src[8] = {f->rho, f->rho_u, f->rho_v, f->rho_E, f->u, f->v, f->p, f->cs};
  for (int k = 0; k < 8; ++k) memcpy(s->a[k], src[k], n * sizeof(double));
}

static int nan_check(const or_config *c, const double *x) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  int found = 0;
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static) reduction(| : found)
#endif
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i)
      found |= isnan(x[CI(i, j)]) ? 1 : 0;
  return found;
}
static int neg_check(const or_config *c, const double *x) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  int found = 0;
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static) reduction(| : found)
#endif
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i)
      found |= (x[CI(i, j)] < 0.0) ? 1 : 0;
  return found;
}

/* L(U) into d[0..3] (time_derivative, mod_fluid.f90:460-516): prims of U get BC-filled in place;
 * the energy source field, when given, is added to d(rho E)/dt (mod_fluid.f90:500-509: a plain
 * array sum, not a field operator). */
static void time_derivative(const or_config *c, const or_geometry *g, stage_buf *U, double **d, const double *src) {
  or_solve(c, g, U->a[0], U->a[4], U->a[5], U->a[6], d[0], d[1], d[2], d[3]);
  if (src) {
    const int ni = c->ni, nj = c->nj, nh = c->nh;
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) d[3][CI(i, j)] = d[3][CI(i, j)] + src[CI(i, j)];
  }
}

/* out = base + L*s (fluid + fluid*real), interior, add-with-flush */
static void combine1(const or_config *c, const stage_buf *base, double **L, double s, stage_buf *out) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  for (int v = 0; v < 4; ++v)
    OR_PAR
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t k = CI(i, j);
        out->a[v][k] = op_add(base->a[v][k], L[v][k] * s);
      }
}
/* out = ((U0*a) + (Uk*b)) + (L*cdt) */
static void combine2(const or_config *c, const stage_buf *U0, double a, const stage_buf *Uk, double b, double **L,
                     double cdt, stage_buf *out) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  for (int v = 0; v < 4; ++v)
    OR_PAR
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) {
        size_t k = CI(i, j);
        double t = op_add(U0->a[v][k] * a, Uk->a[v][k] * b);
        out->a[v][k] = op_add(t, L[v][k] * cdt);
      }
}

static int assign_stage(const or_config *c, stage_buf *s) {
  /* assign_fluid (mod_fluid.f90:718-723) checks for NaNs and recomputes the derived quantities; the
   * reference aborts on a NaN, this restatement reports NANS_FOUND for a NaN in the new stage's
   * rho, u, v or p. */
  or_derived_quantities(c, s->a[0], s->a[1], s->a[2], s->a[3], s->a[4], s->a[5], s->a[6], s->a[7]);
  if (nan_check(c, s->a[0]) || nan_check(c, s->a[4]) || nan_check(c, s->a[5]) || nan_check(c, s->a[6])) return 3;
  return 0;
}

/* NORM2 of (a - b) over the interior, or of a when b is NULL, in array element order (i fastest).
 * NORM2 is a gfortran intrinsic, not reference code: this follows the scaled sum of squares that
 * libgfortran and the gfortran front end use (scale = 1, ssq = 0; a new maximum rescales the sum),
 * which the reference's own norm2 calls inherit; any other order agrees to a few ulps. */
static double norm2_interior(const or_config *c, const double *a, const double *b) {
  const int ni = c->ni, nj = c->nj, nh = c->nh;
  double scale = 1.0, ssq = 0.0;
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i) {
      double x = b ? a[CI(i, j)] - b[CI(i, j)] : a[CI(i, j)];
      if (x != 0.0) {
        double ax = fabs(x);
        if (scale < ax) {
          double r = scale / ax;
          ssq = 1.0 + ssq * r * r;
          scale = ax;
        } else {
          double r = ax / scale;
          ssq = ssq + r * r;
        }
      }
    }
  return scale * sqrt(ssq);
}

/* calculate_convergence (mod_fluid.f90:1162-1205): relative L2 change from first to last; the
 * momentum norms are divided by max(norm2(first), 1). */
static void convergence(const or_config *c, const stage_buf *first, const stage_buf *last, double *out) {
  for (int v = 0; v < 4; ++v) {
    double num = norm2_interior(c, last->a[v], first->a[v]);
    double den = norm2_interior(c, first->a[v], NULL);
    if (v == 1 || v == 2) den = den > 1.0 ? den : 1.0;
    out[v] = num / den;
  }
}

/* Boundary time of each stage, from the flux solver objects the stages call (each solve first adds its dt
 * to the solver's time, slau_solver.fypp:168 and the other solvers alike):
 * - ssp_rk_2_2 / ssp_rk_3_3: stage 1 uses U's own solver, whose time has gained every step's dt since
 *   fluid_t%initialize (the same sum as fluid_t%time); later stages use a stage fluid assigned from an
 *   expression led by U, whose solver restarts from U%time (assign_fluid, mod_fluid.f90:707-711;
 *   add_fluid / fluid_mul_real allocate from the left operand, 617, 635).  All: time + dt.
 * - rk_multistage: stage 1 uses U_prev, a fresh copy of U each step (a local allocatable, 896-898, 952),
 *   whose solver still holds the time U's solver started from, since U%t is never called there: t0 + dt;
 *   later stages U_prev = U_curr (assign_fluid, U_curr led by U_0 = U): time + dt. */
static double stage_bc_time(const or_config *c, const or_fluid *f, double dt, int stage) {
  int multistage = c->integrator != OR_RK_SSP2 && c->integrator != OR_RK_SSP3;
  if (stage == 0) return multistage ? f->t0 + dt : f->t_solver + dt;
  return f->time + dt;
}

int or_integrate(const or_config *c_in, const or_geometry *g, or_fluid *f, double dt) {
  or_config cb = *c_in;
  const or_config *c = &cb;
  size_t n = ncell(c);
  stage_buf U, U1, U2, U3;
  alloc_stage(c, &U);
  alloc_stage(c, &U1);
  alloc_stage(c, &U2);
  alloc_stage(c, &U3);
  double *L[4];
  for (int v = 0; v < 4; ++v) L[v] = (double *)calloc(n, sizeof(double));
  fluid_to_stage(f, &U, n);
  int err = 0;
  stage_buf *fin = NULL;
  if (c->integrator == OR_RK_SSP2 || c->integrator == OR_RK_SSP3) {
    /* ssp_rk_2_2 (mod_fluid.f90:1115-1142) / ssp_rk_3_3 (773-827): the step, its convergence norms, and with
     * allow_subcycle a retry with dt / 10 while any norm exceeds the tolerance (subcycle_check, 856-887).
     * U is untouched until an attempt is kept; every attempt advances U's own solver clock (stage 1). */
    for (;;) {
      err = 0;
      fluid_to_stage(f, &U1, n);
      cb.bc_time = stage_bc_time(c, f, dt, 0);
      f->t_solver = f->t_solver + dt;
      time_derivative(c, g, &U, L, f->src_E);
      combine1(c, &U, L, dt, &U1);
      err |= assign_stage(c, &U1);
      fluid_to_stage(f, &U2, n);
      cb.bc_time = stage_bc_time(c, f, dt, 1);
      time_derivative(c, g, &U1, L, f->src_E);
      if (c->integrator == OR_RK_SSP2) {
        combine2(c, &U, 0.5, &U1, 0.5, L, 0.5 * dt, &U2);
        err |= assign_stage(c, &U2);
        convergence(c, &U1, &U2, f->conv); /* mod_fluid.f90:1132 */
        fin = &U2;
      } else {
        combine2(c, &U, 3.0 / 4.0, &U1, 1.0 / 4.0, L, (1.0 / 4.0) * dt, &U2);
        err |= assign_stage(c, &U2);
        fluid_to_stage(f, &U3, n);
        cb.bc_time = stage_bc_time(c, f, dt, 2);
        time_derivative(c, g, &U2, L, f->src_E);
        combine2(c, &U, 1.0 / 3.0, &U2, 2.0 / 3.0, L, (2.0 / 3.0) * dt, &U3);
        err |= assign_stage(c, &U3);
        convergence(c, &U1, &U3, f->conv); /* mod_fluid.f90:805 */
        fin = &U3;
      }
      int retry = 0;
      if (c->allow_subcycle)
        for (int v = 0; v < 4; ++v) retry |= f->conv[v] > c->subcycle_tolerance;
      if (!retry) break;
      dt = dt / 10.0;
    }
  } else {
    /* rk_multistage (mod_fluid.f90:889-1007) */
    static const double a13[3] = {0.1481, 0.4, 1.0}, a14[4] = {0.0833, 0.2069, 0.4265, 1.0},
                        a15[5] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0}, a23[3] = {0.1918, 0.4929, 1.0},
                        a24[4] = {0.1084, 0.2602, 0.5052, 1.0}, a25[5] = {0.0695, 0.1602, 0.2898, 0.5060, 1.0};
    const double *alpha;
    int ns;
    switch (c->integrator) {
      case OR_RK_1_3: alpha = a13; ns = 3; break;
      case OR_RK_1_4: alpha = a14; ns = 4; break;
      case OR_RK_1_5: alpha = a15; ns = 5; break;
      case OR_RK_2_3: alpha = a23; ns = 3; break;
      case OR_RK_2_4: alpha = a24; ns = 4; break;
      default: alpha = a25; ns = 5; break;
    }
    /* U_0 = U; U_prev starts as a copy of U.  rk_multistage has no subcycling loop (mod_fluid.f90:974-990 is
     * commented out); U's own solver is never called, so its clock stays put. */
    fluid_to_stage(f, &U1, n); /* U_prev */
    fluid_to_stage(f, &U2, n); /* U_curr */
    for (int m = 0; m < ns; ++m) {
      cb.bc_time = stage_bc_time(c, f, dt, m);
      time_derivative(c, g, &U1, L, f->src_E);
      combine1(c, &U, L, dt * alpha[m], &U2);
      err |= assign_stage(c, &U2);
      for (int k = 0; k < 8; ++k) memcpy(U1.a[k], U2.a[k], n * sizeof(double));
    }
    convergence(c, &U, &U2, f->conv); /* mod_fluid.f90:970: U_0 against the last stage */
    fin = &U2;
  }
  /* copy-back of the conserved fields, then sanity check and derived quantities */
  memcpy(f->rho, fin->a[0], n * sizeof(double));
  memcpy(f->rho_u, fin->a[1], n * sizeof(double));
  memcpy(f->rho_v, fin->a[2], n * sizeof(double));
  memcpy(f->rho_E, fin->a[3], n * sizeof(double));
  f->time = f->time + dt; /* ssp_rk_3_3:844, ssp_rk_2_2 and rk_multistage alike */
  f->last_dt = dt;        /* fluid_t%integrate returns the dt it used (mod_fluid.f90:448) */
  int code = 0;
  if (err & 3) code = 3;
  or_derived_quantities(c, f->rho, f->rho_u, f->rho_v, f->rho_E, f->u, f->v, f->p, f->cs);
  if (!code && neg_check(c, f->rho)) code = 1;
  if (!code && neg_check(c, f->p)) code = 2;
  if (!code && (nan_check(c, f->rho) || nan_check(c, f->u) || nan_check(c, f->v) || nan_check(c, f->p))) code = 3;
  free_stage(&U);
  free_stage(&U1);
  free_stage(&U2);
  free_stage(&U3);
  for (int v = 0; v < 4; ++v) free(L[v]);
  return code;
}

long or_run(const or_config *c, const or_geometry *g, or_fluid *f, double max_time, long max_iterations,
            double contour_interval_dt, double *final_time, int *error_code) {
  /* main.f90:97-198 with no I/O and no constant-dt / startup options, starting at the fluid's time
   * (main.f90:100-102 on a restart; 0 otherwise) */
  double time = f->time, next_output_time = time + contour_interval_dt;
  double master_time = time; /* master_puppeteer.f90:206 accumulates dt without the driver's clip */
  long iteration = 0;
  *error_code = 0;
  while (time < max_time && iteration < max_iterations) {
    double dt = or_get_timestep(c, g, f);
    int err = or_integrate(c, g, f, dt);
    dt = f->last_dt; /* master%dt: the dt the step used (a subcycled step's last one) */
    master_time = master_time + dt;
    double new_time = master_time;
    if (err) {
      *error_code = err;
      iteration++;
      time = new_time;
      break;
    }
    if (fabs(time - next_output_time) < EPS) next_output_time = next_output_time + contour_interval_dt;
    if (new_time > next_output_time) {
      double delta_t = fabs(next_output_time - time);
      new_time = time + delta_t;
    }
    time = new_time;
    iteration++;
  }
  *final_time = time;
  return iteration;
}
