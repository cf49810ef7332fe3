/*
 * cato_hip.h -- C ABI of the MI355X-native CATO hot path (libcato_hip.so).
 *
 * The reference (smillerc/fvleg_2d, "CATO") advances the 2D Euler equations
 * through four Fortran plugin interfaces; every entry point below replaces one
 * of them and keeps its argument meaning:
 *
 *   flux_solver_t%solve            src/lib/flux_solvers/flux_solver.f90:66-87   -> cato_residual
 *   boundary_condition_t%apply     src/lib/boundary_conditions/boundary_conditions.f90:60-70 and
 *                                  flux_solver_t%apply_primitive_bc flux_solver.f90:138-185 -> cato_apply_bc
 *   muscl_interpolation_t%interpolate_edge_values
 *                                  src/lib/spatial_reconstruction/muscl_interpolation.f90:37-43 -> cato_interpolate_edges
 *   get_<dir>flux of each solver   e.g. slau_solver.fypp:253-494               -> cato_edge_fluxes
 *   fluid_t%get_timestep           src/lib/fluid/mod_fluid.f90:518-562          -> cato_compute_dt
 *   fluid_t%integrate              src/lib/fluid/mod_fluid.f90:411-450          -> cato_step
 *   program cato time loop         src/main.f90:138-198                         -> cato_run
 *   fluid_t%initialize_from_hdf5   src/lib/fluid/mod_fluid.f90:256-381          -> cato_set_primitive
 *
 * Conventions
 *   - Host arrays are Fortran-ordered (i fastest).  A field argument is a
 *     pointer plus `ld` (elements between consecutive j) and `halo` (number of
 *     ghost layers stored around the interior in that host array), so the
 *     reference's data(ilo-nh:ihi+nh, jlo-nh:jhi+nh) is passed with
 *     halo = nh, ld = ni + 2*nh.
 *   - Every call returns CATO_OK (0) or a negative status; cato_last_error()
 *     describes the last failure of the calling thread.  Nothing aborts.
 *   - The solve's error code of the reference (error.f90:6-9) is returned
 *     through `error_code`: 0 ALL_OK, 1 NEG_DENSITY, 2 NEG_PRESSURE, 3 NANS_FOUND.
 *   - A context owns its device state, one HIP stream and no globals; distinct
 *     contexts may be used from distinct threads.
 */
#ifndef CATO_HIP_H
#define CATO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CATO_HIP_ABI_VERSION 4

/* flux_solver names of the input deck (mod_fluid.f90:223-238) */
enum {
  CATO_FLUX_AUSMPW_PLUS = 0, /* 'AUSMPW+'                                   */
  CATO_FLUX_SLAU = 1,        /* 'SLAU'                                      */
  CATO_FLUX_SLAU2 = 2,       /* 'SLAU2'                                     */
  CATO_FLUX_SD_SLAU = 3,     /* 'SD-SLAU'                                   */
  CATO_FLUX_SD_SLAU2 = 4,    /* 'SD-SLAU2'                                  */
  CATO_FLUX_ROE = 5,         /* 'Roe', 'S-Roe', 'SP-Roe' (reference variant) */
  CATO_FLUX_AUSM_PLUS = 6,   /* 'AUSM+' / 'AUSM+-u'                         */
  CATO_FLUX_HLLC = 7,        /* 'HLLC' (Toro, PVRS wave speeds)             */
  CATO_FLUX_M_AUSMPW_PLUS = 8 /* 'M-AUSMPW+' (m_ausmpw_plus_solver.fypp)   */
};
/* limiter names (muscl_factory.f90:69-90) */
enum {
  CATO_LIMITER_MINMOD = 0,
  CATO_LIMITER_SUPERBEE = 1,
  CATO_LIMITER_VAN_LEER = 2,
  CATO_LIMITER_MLP3 = 3,
  CATO_LIMITER_MLP5 = 4
};
/* boundary condition names (bc_factory.f90:48-64) */
enum {
  CATO_BC_PERIODIC = 0,      /* periodic_bc.fypp (priority 2)                                  */
  CATO_BC_ZERO_GRADIENT = 1, /* zero_gradient_bc.f90 (priority 1)                              */
  CATO_BC_SYMMETRY = 2,      /* symmetry_bc.f90, also 'mirror' / 'reflection' (priority 1)     */
  CATO_BC_OUTFLOW = 3,       /* outlet_bc.f90, 'outflow' / 'outlet' (priority 0): at +x the    */
                             /* characteristic extrapolation to bc_pressure, elsewhere a copy  */
  CATO_BC_PRESSURE_INPUT = 4 /* pressure_input_bc.f90 (priority 0, +x only): inlet/outlet      */
                             /* states of inlet_outlet.f90 for a desired pressure and density  */
};
/* integration_strategy names (mod_fluid.f90:424-446) */
enum {
  CATO_RK_SSP2 = 0, /* 'ssp_rk2'             */
  CATO_RK_SSP3 = 1, /* 'ssp_rk3', 'ssp_rk33' */
  CATO_RK_1_3 = 2,  /* 'rk(1,3)'             */
  CATO_RK_1_4 = 3,
  CATO_RK_1_5 = 4,
  CATO_RK_2_3 = 5,
  CATO_RK_2_4 = 6,
  CATO_RK_2_5 = 7
};
/* arithmetic of reconstruction + edge flux; state, divergence, RK and CFL stay FP64 */
enum { CATO_PRECISION_FP64 = 0, CATO_PRECISION_FP32_FLUX = 1 };

enum {
  CATO_OK = 0,
  CATO_ERR_INVALID = -1,     /* bad argument or configuration */
  CATO_ERR_UNSUPPORTED = -2, /* valid for the reference but not for this path (e.g. a quad grid in a batch) */
  CATO_ERR_HIP = -3,         /* HIP runtime failure */
  CATO_ERR_NO_DEVICE = -4
};

typedef struct cato_config {
  int flux;
  int limiter;
  int integrator;
  int precision;
  int bc_plus_x, bc_minus_x, bc_plus_y, bc_minus_y;
  int n_ghost_layers;     /* 2 (TVD2, MLP3) or 3 (MLP5): mod_input.f90:320-338 */
  int apply_low_mach_fix; /* AUSMPW+ option, deck default true (mod_input.f90:278) */
  double gamma;           /* [physics] polytropic_index */
  double cfl;             /* [time] cfl */
  int tile;               /* -1: library choice; otherwise a tile id from cato_tile_shape */
  double bc_pressure;     /* outflow ambient pressure ([boundary_conditions] outflow_ambient_pressure,
                             mod_input.f90:384-390); must be > 0 when a side is CATO_BC_OUTFLOW.
                             CATO_BC_PRESSURE_INPUT: constant_bc_pressure_value (mod_input.f90:375) */
  double bc_density;      /* CATO_BC_PRESSURE_INPUT: bc_density (mod_input.f90:372), used with the
                             constant pressure; both must be >= tiny when a stage uses them
                             (get_desired_pressure / _density, pressure_input_bc.f90:204-229) */
} cato_config;

typedef struct cato_geom {
  int ni, nj;             /* interior cells */
  const double *node_x;   /* (ni+2nh+1) x (nj+2nh+1) node coordinates incl. ghost nodes,   */
  const double *node_y;   /* column-major, as the pycato deck stores /x and /y            */
} cato_geom;

typedef struct cato_ctx cato_ctx;

const char *cato_last_error(void);
int cato_abi_version(void);

/* Tile shapes of the fused stage kernel (used by the autotuner). */
int cato_num_tiles(void);
int cato_tile_shape(int tile, int *tile_x, int *tile_y);

/* fluid_t%initialize + grid_block_2d_t geometry: validates the configuration,
 * builds the per-cell geometry from the nodes (quad_cell.f90:69-229) and
 * allocates the device-resident state.  Rectilinear grids keep separable 1D
 * geometry; any other quadrilateral grid keeps per-cell volumes, edge lengths,
 * min_dx / min_dy and edge normals (the flux path of flux_solver.f90:254-367,
 * 398-422 reads them) and runs the 16 x 16 tile.  A non-positive cell volume is
 * CATO_ERR_INVALID. */
int cato_ctx_create(const cato_config *cfg, const cato_geom *geom, cato_ctx **out);
int cato_ctx_destroy(cato_ctx *ctx);

/* Initial / restart state from primitives (mod_fluid.f90:339-363): conserved
 * variables are formed on the device; the first stage uses these primitives. */
int cato_set_primitive(cato_ctx *ctx, const double *rho, const double *u, const double *v, const double *p, int ld,
                       int halo);
/* State from conserved variables (primitives are recovered with the Mach filter). */
int cato_set_conserved(cato_ctx *ctx, const double *rho, const double *rho_u, const double *rho_v,
                       const double *rho_E, int ld, int halo);
int cato_get_conserved(cato_ctx *ctx, double *rho, double *rho_u, double *rho_v, double *rho_E, int ld, int halo);
/* u, v, p, cs of fluid_t after calculate_derived_quantities (any pointer may be NULL). */
int cato_get_primitive(cato_ctx *ctx, double *rho, double *u, double *v, double *p, double *cs, int ld, int halo);

/* fluid_t%get_timestep */
int cato_compute_dt(cato_ctx *ctx, double *dt);
/* Allocation-free variant of cato_compute_dt, with no counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The CFL time step (fluid_t%get_timestep, as cato_compute_dt computes it) of
 * a state the caller holds in device memory: rho, u, v, p are device arrays in the host API's layout
 * (ld, halo) of the context's grid; dt is written to the device double *dt_device on the context's
 * stream, with no host synchronisation, no allocation and no memset.  The per-workgroup partial
 * minima live in one fixed library array per GPU (4096 partials, about 48 KiB, held until the
 * process ends); a grid with more than 4096 x 256 cells streams through 4096 workgroups, so no grid
 * is too large.  Calls that overlap on one GPU (other contexts or streams) share that array and
 * race: serialise them. */
int cato_compute_dt_noalloc(cato_ctx *ctx, const double *rho, const double *u, const double *v, const double *p,
                            int ld, int halo, double *dt_device);
/* fluid_t%integrate with a host-given dt (one full SSP/RK step). */
int cato_step(cato_ctx *ctx, double dt, int *error_code);
/* main.f90 time loop on the device: while time < max_time and iteration <
 * max_iterations, dt = CFL (or constant_dt > 0), integrate, advance the driver
 * time with the contour-interval clip.  The driver clock starts at the fluid time
 * (main.f90:100-102; 0 unless cato_set_time or earlier steps moved it).  Returns
 * iterations, driver time and the first error code.  The host only polls every
 * few steps. */
int cato_run(cato_ctx *ctx, double max_time, long long max_iterations, double contour_interval_dt, double constant_dt,
             long long *iterations, double *time, int *error_code);
/* The other dt choices of the main.f90 time loop (main.f90:140-158) for the
 * following cato_run calls: with enable_startup, dt = startup_dt while the driver
 * time <= startup_period and the CFL dt after it ([time] enable_startup_time,
 * startup_period, startup_delta_t); otherwise initial_dt > 0 is the dt of the
 * driver's iteration 0 ([time] initial_delta_t without use_constant_delta_t),
 * where the driver's iteration is start_iteration (a restart's iteration) plus
 * the steps of this cato_run.  All in code units; 0 / 0 / 0 / 0 / 0 (the default)
 * leaves CFL or constant_dt.  exact_landing != 0 (an option outside the reference,
 * for runs that need the state at a given time) also shortens a step that would
 * carry the fluid time past the next contour time or max_time, so the state lands
 * on them; the reference (exact_landing = 0) clips only the driver time
 * (main.f90:178-194).  cato_advance always uses the CFL dt. */
int cato_set_loop_dt(cato_ctx *ctx, double initial_dt, int enable_startup, double startup_period, double startup_dt,
                     long long start_iteration, int exact_landing);
/* Energy source of fluid_t%time_derivative (mod_fluid.f90:500-509, with the field
 * that source_t%integrate returns for the step, ssp_rk_3_3 mod_fluid.f90:767-770):
 * q_dot (interior ni x nj, host layout ld/halo) is added to d(rho E)/dt in every
 * stage of every following cato_step / cato_run / cato_advance until it is
 * replaced or cleared (q_dot = NULL).  cato_residual (solver%solve) never adds it. */
int cato_set_energy_source(cato_ctx *ctx, const double *q_dot, int ld, int halo);
/* pressure_input_bc_t%read_pressure_input (pressure_input_bc.f90:93-178) after the file is read:
 * n >= 2 rows of time (strictly increasing), pressure and density, already scaled
 * (bc_pressure_scale_factor) and in code units.  Every following stage evaluates them
 * by linear interpolation at its boundary time (finterp's linear_interp_1d, extrapolating
 * beyond the ends), replacing the constant bc_pressure / bc_density; n = 0 returns to
 * the constants.  Errors as initialize_1d: n < 2 or times not increasing. */
int cato_set_pressure_input(cato_ctx *ctx, const double *time, const double *pressure, const double *density, int n);
/* fluid_t%set_time + the time the flux solver starts from (fluid_t%initialize,
 * mod_fluid.f90:169,240; master_puppeteer%set_time): the fluid time every following
 * step advances, the start of cato_run's driver clock, and the origin of the boundary
 * times of time-dependent boundaries.  cato_set_primitive / cato_set_conserved return
 * the fluid to this time (default 0). */
int cato_set_time(cato_ctx *ctx, double time);
/* Enqueue exactly n steps with CFL dt and no host synchronisation (benchmarks). */
int cato_advance(cato_ctx *ctx, long long n);
/* The device-side loop of the last cato_run / cato_advance, after waiting for the context's stream:
 * steps taken, driver time and the first error code (0 ALL_OK, 1, 2, 3 as error.f90:6-9).  A step
 * after an error, or past the loop's end, returns at once without work, so a caller of cato_advance
 * checks here that all n steps ran.  No counterpart in the reference (its loop is on the host). */
int cato_loop_status(cato_ctx *ctx, long long *iterations, double *time, int *error_code);
/* Device-side loop of cato_run / cato_advance as a captured hipGraph of 32 steps
 * (enable != 0, the default) or as individual kernel launches.  Both give
 * identical results; timing (cato_set_timing) always launches individually.
 * No counterpart in the reference: its loop is host-driven (main.f90:138-198). */
int cato_set_device_graph(cato_ctx *ctx, int enable);
int cato_synchronize(cato_ctx *ctx);

/* Step bookkeeping of fluid_t%integrate: calculate_convergence (mod_fluid.f90:1162-1205),
 * the relative L2 change [rho, rho u, rho v, rho E] between the last stage and U_1
 * (SSP-RK2/RK3, mod_fluid.f90:805,1132) or U_0 (rk(k,n), mod_fluid.f90:970), with
 * the momenta divided by max(norm2(first), 1); and write_residual_history
 * (mod_fluid.f90:1207-1223).  Off by default.  While on, the last stage of every step
 * also reads the compared state and leaves per-workgroup sums that the step's
 * finalize reduces in a fixed order (norms agree with NORM2 to ~1e-15 relative).
 * Enabling (or disabling) clears the collected rows; the row counters restart at
 * 0 here and at cato_set_primitive / cato_set_conserved. */
int cato_set_diagnostics(cato_ctx *ctx, int enable);
/* The four norms of the last step (error if diagnostics are off). */
int cato_get_convergence(cato_ctx *ctx, double norms[4]);
/* Moves up to max_rows collected rows (oldest first) into rows, 7 doubles each:
 * iteration, time, dt, norms[4] -- the columns of residual_hist.csv.  cato_step and
 * cato_run record one row per step; cato_advance records none.  rows = NULL only
 * reports the number of rows waiting in *n_rows. */
int cato_get_history(cato_ctx *ctx, double *rows, long long max_rows, long long *n_rows);

/* [time] allow_subcycle / subcycle_error_tolerance (ssp_rk_2_2 / ssp_rk_3_3 with
 * subcycle_check, mod_fluid.f90:816-823,856-887; deck keys mod_input.f90:255-256):
 * while any convergence norm of a cato_step exceeds the tolerance, the step is
 * recomputed from the same state with dt / 10.  SSP-RK2/3 only (rk_multistage has
 * no retry loop); host-driven steps only (cato_run / cato_advance refuse an SSP
 * context with it on).  cato_last_dt: the dt the last cato_step used (fluid_t%dt). */
int cato_set_subcycle(cato_ctx *ctx, int enable, double tolerance);
int cato_last_dt(cato_ctx *ctx, double *dt);
/* Per-launch timing of the fused stage kernel, measured with HIP events on the
 * context's stream while enabled. */
int cato_set_timing(cato_ctx *ctx, int enable);
int cato_get_timing(cato_ctx *ctx, double *stage_ms_total, long long *stage_launches, double *other_ms_total,
                    long long *other_launches);

/* ---- operator-level entry points (arrays carry n_ghost_layers halos, ld >= ni+2nh) ---- */
/* apply_primitive_bc: fills the ghost cells of rho, u, v, p in place. */
int cato_apply_bc(cato_ctx *ctx, double *rho, double *u, double *v, double *p, int ld);
/* interpolate_edge_values for one halo'd field q: i_edges and j_edges are
 * (2, 0:ni, 0:nj) arrays of (L, R) states. */
int cato_interpolate_edges(cato_ctx *ctx, const double *q, int ld, double *i_edges, double *j_edges);
/* get_<dir>flux: edges = 8 consecutive (2, 0:ni, 0:nj) arrays (rho_i, rho_j, u_i, u_j,
 * v_i, v_j, p_i, p_j), for M-AUSMPW+ followed by the same 8 of the superbee reconstruction;
 * p = halo'd cell pressure; iflux (4, 0:ni, 1:nj), jflux (4, 1:ni, 0:nj). */
int cato_edge_fluxes(cato_ctx *ctx, const double *edges, const double *p, int ld, double *iflux, double *jflux);
/* flux_solver_t%solve: BCs are applied to rho, u, v, p in place (intent(inout)),
 * d_* receive d(rho, rho u, rho v, rho E)/dt with zeroed halos. */
int cato_residual(cato_ctx *ctx, double *rho, double *u, double *v, double *p, double *d_rho, double *d_rho_u,
                  double *d_rho_v, double *d_rho_E, int ld);

/* ---- batched form: many independent solves in one persistent launch per stage ---- */
typedef struct cato_batch cato_batch;
/* Each solve k has its own config and (rectilinear) geometry; all solves share the precision and tile. */
int cato_batch_create(int n, const cato_config *cfgs, const cato_geom *geoms, int precision, int tile,
                      cato_batch **out);
int cato_batch_destroy(cato_batch *b);
int cato_batch_set_primitive(cato_batch *b, int k, const double *rho, const double *u, const double *v,
                             const double *p, int ld, int halo);
int cato_batch_get_conserved(cato_batch *b, int k, double *rho, double *rho_u, double *rho_v, double *rho_E, int ld,
                             int halo);
/* Time-tabled +x pressure input of solve k (as cato_set_pressure_input, pressure_input_bc.f90:93-178); a
 * solve with CATO_BC_PRESSURE_INPUT and no table uses its config's constant bc_pressure / bc_density.  Its
 * column pass runs before every stage of the batch (the served form takes the constant drive). */
int cato_batch_set_pressure_input(cato_batch *b, int k, const double *time, const double *pressure,
                                  const double *density, int n);
/* Run every solve to its own max_time (main.f90 semantics per solve); steps are
 * replayed from a captured hipGraph.  iterations/times/errors may be NULL. */
int cato_batch_run(cato_batch *b, const double *max_time, const double *contour_interval_dt, long long max_iterations,
                   int use_graph, long long *iterations, double *times, int *errors);

/* Device memory one solve of this configuration and grid occupies (its HBM
 * footprint, as the serving form charges it against its budget).  Needs no
 * device. */
int cato_device_bytes(const cato_config *cfg, const cato_geom *geom, unsigned long long *bytes);

/* ---- serving form: continuous batching under an HBM budget ----
 * No counterpart in the reference (one solve per program run there).  A
 * server owns one device arena of hbm_budget_bytes, allocated at create; the
 * request loop itself never allocates device memory.  All requests of a
 * server share its precision and the integrator of the first request; a
 * request with CATO_BC_PRESSURE_INPUT takes its config's constant drive. */
typedef struct cato_server cato_server;
typedef struct cato_request {
  cato_config cfg;
  cato_geom geom;
  const double *rho, *u, *v, *p; /* interior primitives, Fortran order, ld = ni */
  double max_time;               /* main.f90 loop semantics per request */
  double contour_interval_dt;    /* 0: max_time */
  double *out_rho, *out_rho_u, *out_rho_v, *out_rho_E; /* interior conserved result (ld = ni); may be NULL */
} cato_request;
int cato_server_create(double hbm_budget_bytes, int max_slots, int precision, int steps_per_tick,
                       cato_server **out);
int cato_server_destroy(cato_server *s);
/* Queue a request (inputs are copied; the output pointers must stay valid
 * until the request completes).  Returns the request id (>= 0) or a status. */
int cato_server_submit(cato_server *s, const cato_request *req, double arrival_time);
/* One tick: retire finished solves (results copied out, memory returned),
 * admit queued requests in arrival order while a slot and arena memory are
 * free, then advance every resident solve by steps_per_tick steps with one
 * persistent launch per RK stage.  Any output pointer may be NULL. */
int cato_server_tick(cato_server *s, int *n_resident, int *n_queued, int *n_completed_total);
/* Completion record of request `id`: done flag, completion time on the
 * server clock, iterations and solve error code (-1: request rejected). */
int cato_server_poll(cato_server *s, int id, int *done, double *completion_time, long long *iterations,
                     int *error_code);
/* Seconds since cato_server_create. */
double cato_server_clock(cato_server *s);

/* ---- one solve split into px x py sub-domains on one GPU ----
 * The shape a run of one large solve over several GPUs takes: the reference's near-square blocks per
 * coarray image (mod_parallel.f90:41-124) with nh-deep edge and corner halos of the primitives exchanged
 * after every RK stage (field_2d.f90:551-654) and the CFL dt as the minimum over the blocks
 * (mod_fluid.f90:560).  Here every block is a sub-domain on the same GPU and the exchange is a pack kernel
 * (each block gathers what its neighbours need into a send buffer) and an unpack kernel (each block pulls
 * its neighbours' send regions into its halo store).  Results are bit-identical to the unsplit solve.
 * FP64 form, rectilinear grids; periodic, zero-gradient and symmetry sides.  Host arrays as for cato_ctx
 * (whole grid). */
typedef struct cato_split cato_split;
int cato_split_create(const cato_config *cfg, const cato_geom *geom, int px, int py, cato_split **out);
int cato_split_destroy(cato_split *split);
int cato_split_set_primitive(cato_split *split, const double *rho, const double *u, const double *v,
                             const double *p, int ld, int halo);
int cato_split_get_conserved(cato_split *split, double *rho, double *rho_u, double *rho_v, double *rho_E, int ld,
                             int halo);
int cato_split_compute_dt(cato_split *split, double *dt);
int cato_split_step(cato_split *split, double dt, int *error_code);
/* GPU time of the halo exchanges so far (pack + unpack of every block, HIP events) and their number. */
int cato_split_exchange_time(cato_split *split, double *ms_total, long long *exchanges);

/* ---- measurement utility (no counterpart in the reference) ----
 * The device's measured HBM ceiling: kind 0 streams a copy b -> a, kind 1 a triad a = b + s c, over
 * FP64 arrays of bytes_per_array each (allocated and freed inside), reps timed launches; returns the
 * median rate in GB/s of compulsory traffic (2 or 3 arrays) and its launch time. */
int cato_hbm_probe(int kind, unsigned long long bytes_per_array, int reps, double *gb_per_s, double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
