/*
 * CATO hot-path oracle: a plain-C restatement of the reference solver's
 * per-timestep arithmetic, used ONLY as the checker by tests/, by
 * __graft_entry__.smoke() and as the CPU-baseline leg of bench.py.
 * It is never linked into, called by, or shipped with the HIP product path.
 *
 * Layout follows the reference: every cell field is a halo'd 2D array stored
 * column-major (i fastest), cell index i in [1-nh, ni+nh], j in [1-nh, nj+nh]
 * (src/lib/field/field_2d.f90:477-483).  Edge states are (L/R, i, j) arrays,
 * edge fluxes are (4, i, j) arrays (src/lib/flux_solvers/slau_solver.fypp:209-210).
 *
 * Parity status per function is stated in oracle/README.md.  The HLLC flux follows textbook Toro
 * (the reference's body is incomplete) and is pinned against what that scheme fixes.  Parity
 * unpinned (no check independent of this restatement exists): the boundary time each stage of
 * or_integrate evaluates a time-dependent pressure input at (stage_bc_time, read from the
 * reference's solver clocks).
 */
#ifndef CATO_ORACLE_H
#define CATO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Flux solver ids (the reference's input%flux_solver names, mod_fluid.f90:223-238). */
enum {
  OR_FLUX_AUSMPW_PLUS = 0, /* 'AUSMPW+'                 ausmpw_plus_solver.fypp */
  OR_FLUX_SLAU = 1,        /* 'SLAU'                    slau_solver.fypp        */
  OR_FLUX_SLAU2 = 2,       /* 'SLAU2'                                           */
  OR_FLUX_SD_SLAU = 3,     /* 'SD-SLAU'                                         */
  OR_FLUX_SD_SLAU2 = 4,    /* 'SD-SLAU2'                                        */
  OR_FLUX_ROE = 5,         /* 'Roe' / 'S-Roe' / 'SP-Roe' roe_solver.fypp        */
  OR_FLUX_AUSM_PLUS = 6,   /* 'AUSM+' (= AUSM+-u flags)  ausm_plus_solver.f90    */
  OR_FLUX_HLLC = 7,        /* 'HLLC' (textbook Toro; reference body incomplete) */
  OR_FLUX_M_AUSMPW_PLUS = 8 /* 'M-AUSMPW+'                m_ausmpw_plus_solver.fypp */
};

/* Reconstruction ids (muscl_factory.f90:69-90). */
enum { OR_LIM_MINMOD = 0, OR_LIM_SUPERBEE = 1, OR_LIM_VAN_LEER = 2, OR_LIM_MLP3 = 3, OR_LIM_MLP5 = 4 };

/* Boundary condition ids (bc_factory.f90:48-64). */
enum { OR_BC_PERIODIC = 0, OR_BC_ZERO_GRADIENT = 1, OR_BC_SYMMETRY = 2, OR_BC_OUTFLOW = 3, OR_BC_PRESSURE_INPUT = 4 };

/* Time integrators (mod_fluid.f90:424-446). */
enum {
  OR_RK_SSP2 = 0, OR_RK_SSP3 = 1,
  OR_RK_1_3 = 2, OR_RK_1_4 = 3, OR_RK_1_5 = 4, OR_RK_2_3 = 5, OR_RK_2_4 = 6, OR_RK_2_5 = 7
};

typedef struct {
  int ni, nj, nh;
  int flux, limiter, integrator;
  int bc_plus_x, bc_minus_x, bc_plus_y, bc_minus_y;
  int apply_low_mach_fix;
  double gamma;
  double cfl;
  double bc_pressure; /* outflow ambient pressure (outlet_bc.f90:55); pressure_input constant pressure */
  /* pressure_input (+x, pressure_input_bc.f90): constant density (apply_constant_bc_pressure), or, with
   * pi_n >= 2, pressure and density linear in time over the table (pi_t strictly increasing); bc_time is
   * the time the boundary is evaluated at (boundary_condition_t%time, set per stage by or_integrate). */
  double bc_density;
  int pi_n;
  const double *pi_t, *pi_p, *pi_rho;
  double bc_time;
  /* [time] allow_subcycle / subcycle_error_tolerance (mod_input.f90:255-256, mod_fluid.f90:171-172) */
  int allow_subcycle;
  double subcycle_tolerance;
} or_config;

/* Per-cell geometry of the halo'd block (grid_block_2d.f90:352-391). */
typedef struct {
  int ni, nj, nh;
  double *volume;  /* (ncx, ncy)        */
  double *dx, *dy; /* (ncx, ncy)        */
  double *edge_lengths; /* (4, ncx, ncy)    */
  double *edge_norm;    /* (2, 4, ncx, ncy) */
} or_geometry;

/* Number of halo'd cells in i and j. */
int or_ncx(int ni, int nh);
int or_ncy(int nj, int nh);

/* quad_cell.f90:69-229 applied to every cell of the halo'd block; nodes are
 * (ni+2nh+1, nj+2nh+1) column-major including ghost nodes. Returns 0 or -1
 * on a non-positive volume. Arrays in g must be preallocated by the caller. */
int or_geometry_from_nodes(int ni, int nj, int nh, const double *node_x, const double *node_y, or_geometry *g);

/* equation_of_state.f90:297-324 (initial prim -> conserved, interior only). */
void or_prim_to_cons(const or_config *c, const double *rho, const double *u, const double *v, const double *p,
                     double *rho_u, double *rho_v, double *rho_E);
/* mod_fluid.f90:564-591 (cons -> prim, cs, Mach filter; interior only). */
void or_derived_quantities(const or_config *c, const double *rho, const double *rho_u, const double *rho_v,
                           const double *rho_E, double *u, double *v, double *p, double *cs);
/* equation_of_state.f90:241-253 (cs only; used at initialization). */
void or_sound_speed(const or_config *c, const double *rho, const double *p, double *cs);

/* flux_solver.f90:138-185 with periodic_bc.fypp / zero_gradient_bc.f90 / symmetry_bc.f90 on one image. */
void or_apply_bcs(const or_config *c, double *rho, double *u, double *v, double *p);

/* linear_interp_1d%evaluate (src/third_party/finterp/linear_interpolation_module.f90:730-762). */
double or_interp_linear(int n, const double *t, const double *f, double x);

/* One boundary_condition_t%apply call; loc 0:+x 1:-x 2:+y 3:-y. */
void or_apply_bc_single(const or_config *c, int type, int loc, double *rho, double *u, double *v, double *p);

/* muscl_tvd2.fypp:126-227 / muscl_mlp.fypp:117-284. Edge arrays are (2, 0:ni, 0:nj). */
void or_interpolate_edges(const or_config *c, const double *q, double *i_edges, double *j_edges);

/* Edge fluxes: iflux (4, 0:ni, 1:nj), jflux (4, 1:ni, 0:nj). */
void or_edge_fluxes(const or_config *c, const or_geometry *g, const double *rho, const double *u, const double *v,
                    const double *p, const double *edges /* 8 arrays: rho_i,rho_j,u_i,u_j,v_i,v_j,p_i,p_j */,
                    double *iflux, double *jflux);

/* The same with the superbee edge states M-AUSMPW+ blends in (8 more arrays in the same order;
 * NULL for the other fluxes).  or_edge_fluxes takes them from edges + 8 arrays for M-AUSMPW+. */
void or_edge_fluxes_sb(const or_config *c, const or_geometry *g, const double *rho, const double *u, const double *v,
                       const double *p, const double *edges, const double *edges_sb, double *iflux, double *jflux);

/* flux_solver.f90:187-459. d_dt arrays are halo'd; halo zeroed. */
void or_flux_split_edges(const or_config *c, const or_geometry *g, const double *iflux, const double *jflux,
                         double *d_rho, double *d_rho_u, double *d_rho_v, double *d_rho_E);

/* flux_solver_t%solve: BCs (in place on rho,u,v,p) + reconstruction + fluxes + divergence. */
void or_solve(const or_config *c, const or_geometry *g, double *rho, double *u, double *v, double *p,
              double *d_rho, double *d_rho_u, double *d_rho_v, double *d_rho_E);

/* Fluid state: 4 conserved + u,v,p,cs (all halo'd); src_E: optional energy source field
 * (halo'd, NULL for none) added to d(rho E)/dt in every stage (mod_fluid.f90:500-509);
 * conv: the L2-norm convergence of the last or_integrate (rho, rho u, rho v, rho E;
 * calculate_convergence, mod_fluid.f90:1162-1205). */
typedef struct {
  double *rho, *rho_u, *rho_v, *rho_E;
  double *u, *v, *p, *cs;
  const double *src_E;
  double conv[4];
  /* fluid_t%time (advanced by dt per or_integrate) and the time its flux solver started from
   * (fluid_t%initialize, mod_fluid.f90:169,240): the boundary times of a step derive from both */
  double time, t0;
  /* the time of the flux solver the first stage of ssp_rk_2_2 / ssp_rk_3_3 calls (U's own): it gains the
   * dt of every attempt, retried ones included; and the dt the last or_integrate used */
  double t_solver, last_dt;
} or_fluid;

/* fluid_t%initialize from primitive interior values (mod_fluid.f90:248-252, 339-363). */
void or_fluid_init_from_prims(const or_config *c, or_fluid *f);
/* fluid_t%get_timestep (mod_fluid.f90:518-562). */
double or_get_timestep(const or_config *c, const or_geometry *g, const or_fluid *f);
/* fluid_t%integrate for one step (ssp_rk_2_2 / ssp_rk_3_3 / rk_multistage). Returns error code
 * (0 ok, 1 negative density, 2 negative pressure, 3 NaN found) as the library does. */
int or_integrate(const or_config *c, const or_geometry *g, or_fluid *f, double dt);

/* main.f90:117-198 time loop (no contour I/O). Runs until time >= max_time or max_iterations.
 * contour_interval_dt drives the dt clip quirk of main.f90:188-194. Returns number of steps. */
long or_run(const or_config *c, const or_geometry *g, or_fluid *f, double max_time, long max_iterations,
            double contour_interval_dt, double *final_time, int *error_code);

#ifdef __cplusplus
}
#endif
#endif
