// Shared enumerations and launch-argument structures of the CATO HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cato_hip.h"

namespace cato {

enum FluxId {
  FLUX_AUSMPW_PLUS = CATO_FLUX_AUSMPW_PLUS,
  FLUX_SLAU = CATO_FLUX_SLAU,
  FLUX_SLAU2 = CATO_FLUX_SLAU2,
  FLUX_SD_SLAU = CATO_FLUX_SD_SLAU,
  FLUX_SD_SLAU2 = CATO_FLUX_SD_SLAU2,
  FLUX_ROE = CATO_FLUX_ROE,
  FLUX_AUSM_PLUS = CATO_FLUX_AUSM_PLUS,
  FLUX_HLLC = CATO_FLUX_HLLC,
  FLUX_M_AUSMPW_PLUS = CATO_FLUX_M_AUSMPW_PLUS,
  FLUX_COUNT = 9
};

enum LimId {
  LIM_MINMOD = CATO_LIMITER_MINMOD,
  LIM_SUPERBEE = CATO_LIMITER_SUPERBEE,
  LIM_VAN_LEER = CATO_LIMITER_VAN_LEER,
  LIM_MLP3 = CATO_LIMITER_MLP3,
  LIM_MLP5 = CATO_LIMITER_MLP5,
  LIM_COUNT = 5
};

enum BcId {
  BC_PERIODIC = CATO_BC_PERIODIC,
  BC_ZERO_GRADIENT = CATO_BC_ZERO_GRADIENT,
  BC_SYMMETRY = CATO_BC_SYMMETRY,
  BC_OUTFLOW = CATO_BC_OUTFLOW,
  BC_PRESSURE_INPUT = CATO_BC_PRESSURE_INPUT
};

// +x sides whose ghost layers all read one stored column ni+1 of the primitive
// arrays (filled per stage by the producer of the primitives).
__host__ __device__ inline bool ghost_column_bc(int bc) { return bc == BC_OUTFLOW || bc == BC_PRESSURE_INPUT; }

// How a stage kernel combines the residual L(U_k) with the stored states.
enum StageMode {
  MODE_FIRST = 0,    // out = U_k + L*(c*dt)                 (SSP stage 1; c = 1)
  MODE_COMBINE = 1,  // out = ((U_0*a) + (U_k*b)) + L*(c*dt) (SSP stages >= 2)
  MODE_BASE = 2,     // out = U_0 + L*(dt*c)                 (rk(k,n) low-storage stages)
  MODE_RESIDUAL = 3  // out = L                              (flux_solver_t%solve)
};

// Device-resident scalars of one solve's time loop (main.f90:138-198).
struct LoopState {
  double dt;                // dt of the next step
  double time;              // driver time (with the contour-interval clip)
  double master_time;       // unclipped accumulated time (master_puppeteer.f90:206)
  double next_output_time;  // contour cadence that drives the clip
  double contour_dt;
  double max_time;
  double constant_dt;       // > 0: use_constant_delta_t
  long long iteration;
  long long max_iterations;
  int done;
  int error;
  int pad[2];
  // main.f90:140-158 dt choices besides constant_dt: a startup period of
  // startup_dt while time <= startup_period (startup != 0), else initial_dt on
  // the driver's iteration 0 (iter0 + iteration == 0; iter0 = the iteration
  // count the run resumes from).  exact != 0: the dt is also clipped so that
  // the state's own time lands on the next contour time and on max_time (an
  // option outside the reference, whose clip moves only the driver time)
  double initial_dt;
  double startup_period;
  double startup_dt;
  long long iter0;
  int startup;
  int exact;
  int pad2[2];
};

// Convergence diagnostics of the step a finalize closes (calculate_convergence,
// mod_fluid.f90:1162-1205): the per-block sums the last stage wrote, the four
// norms, and an optional device ring of history rows for the device-side loop.
constexpr int kHistRow = 7;  // iteration, time, dt, L2 norms of rho, rho u, rho v, rho E
struct ConvOut {
  const double *sums;  // 8 arrays of nb block sums (null: diagnostics off)
  double *norms;       // 4 norms of the last step
  double *hist;        // hist_cap rows of kHistRow, row k of a run at k % hist_cap (null: none)
  int hist_cap;
};

// Per-field pointers of one conserved state, interior only, pitched rows.
struct State4 {
  double *f[4];
};

struct StageArgs {
  int ni, nj, nh, pitch;     // interior cells, halo width, row pitch (elements)
  int mode;
  int check_input;           // 1: flag NaN in the input primitives (the caller's, first stage)
  int last_stage;            // 1: write CFL partials and sanity flags for the new state
  int low_mach_fix;
  int fast_ops;              // 1: FastOps edge arithmetic allowed on in-range tiles (gamma in (1, 4])
  int bc[4];                 // +x, -x, +y, -y
  int tiles_x, tiles_y;      // tile grid
  double gamma, cfl;
  double gas[3];             // gas_constants(gamma): the fluxes' gamma / (gamma - 1), ...
  double bc_pressure;        // outflow ambient pressure (BC_OUTFLOW at +x)
  double a, b, c;            // stage coefficients
  const double *uk[4];       // U_k at the cell (MODE_FIRST / MODE_COMBINE)
  const double *u0[4];       // U_0 (MODE_COMBINE / MODE_BASE)
  const double *prim[4];     // rho, u, v, p of U_k (read with halo through the BC map)
  double *out[4];            // U_{k+1}
  double *prim_out[4];       // rho, u, v, p of U_{k+1} (null in MODE_RESIDUAL)
  const double *src_E;       // energy source added to d(rho E)/dt (null: none; always null in MODE_RESIDUAL)
  const double *dxi, *lxi;   // per halo'd i: x spacing and x-edge length
  const double *dyj, *lyj;   // per halo'd j
  // Uniform rectilinear grid (every interior cell alike): the cell geometry the stage epilogue
  // derives from dxi..lyj, computed once on the host with the same operations: dx, dy, the
  // bottom / right / top / left edge lengths after their equalisation, the volume and, when
  // the volume is a power of two (ugeom_pow2), its exact reciprocal.  ugeom = 0: per cell
  int ugeom, ugeom_pow2;
  double ug[8];
  const LoopState *loop;     // dt source and done flag (may be null -> dt_value)
  double dt_value;
  double *block_min;         // per-block CFL minimum (last stage)
  int *block_err;            // per-block sanity flags
  const double *conv_first[4];  // last stage: the state calculate_convergence compares against (null: off)
  double *block_conv;        // per-block sums for the convergence norms, 8 arrays of tiles_x*tiles_y
  // A sub-domain of a split solve: sides whose bit is set (1 +x, 2 -x, 4 +y, 8 -y) take their ghost
  // cells from the halo store (neighbour primitives exchanged after every stage) instead of a boundary
  // condition; layout in halo_store_index
  int halo_mask;
  const double *halo;
  // General quadrilateral grid (null: rectilinear, dxi..lyj): Q_COUNT per-cell arrays of quad_cell
  // geometry over the halo'd cells, quad_ncx = ni + 2 nh per row, quad_n = quad_ncx (nj + 2 nh) per array
  const double *quad;
  int quad_ncx;
  long long quad_n;
  // Elements between the four variables of every state and primitive set above: the kernel reads
  // uk[0] + v * fstride (likewise u0, prim, out, prim_out, conv_first), so it keeps one base pointer
  // per set in scalar registers instead of four (validate_stage_args checks the layout)
  long long fstride;
};

// Per-cell geometry of a general quadrilateral grid (quad_cell.f90:69-229), one
// array each over the halo'd cells, column-major: the edge lengths of the
// bottom, right, top and left edges, the volume, min_dx / min_dy, and the unit
// normals of the right edge (the i-flux) and the top edge (the j-flux).
enum QuadSlot {
  Q_LEN0 = 0, Q_LEN1, Q_LEN2, Q_LEN3, Q_VOL, Q_MDX, Q_MDY, Q_NRX, Q_NRY, Q_NTX, Q_NTY, Q_COUNT
};

// Halo store of a sub-domain: the +x and -x strips [var][layer][row] (4 nh nj each), the +y and -y strips
// [var][layer][col] (4 nh ni each) and the corners +x+y, -x+y, +x-y, -x-y [var][layer_y][layer_x]
// (4 nh^2 each).  Layer k of a +side is ghost n + 1 + k, of a -side ghost k - nh + 1.
__host__ __device__ inline size_t halo_store_size(int ni, int nj, int nh) {
  return (size_t)8 * nh * (ni + nj) + (size_t)16 * nh * nh;
}
__host__ __device__ inline size_t halo_region_offset(int region, int ni, int nj, int nh) {
  const size_t xs = (size_t)4 * nh * nj, ys = (size_t)4 * nh * ni, cs = (size_t)4 * nh * nh;
  if (region < 2) return region * xs;
  if (region < 4) return 2 * xs + (region - 2) * ys;
  return 2 * xs + 2 * ys + (region - 4) * cs;
}
__host__ __device__ inline size_t halo_region_size(int region, int ni, int nj, int nh) {
  return region < 2 ? (size_t)4 * nh * nj : region < 4 ? (size_t)4 * nh * ni : (size_t)4 * nh * nh;
}

}  // namespace cato
