// Internal host-side types of the CATO HIP runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "aux_kernels.hpp"
#include "stage_dispatch.hpp"

namespace cato {

int fail(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);

// Separable geometry of a rectilinear block, indexed by halo'd cell (0-based):
// dxi = x(i+1)-x(i), lxi = sqrt(dxi**2 + 0**2) (bottom/top edge length), same for j.
// Any other grid (quad = true): the quad_cell geometry of every halo'd cell in
// Q_COUNT arrays of (ni + 2 nh) (nj + 2 nh) values (QuadSlot order); the 1D
// arrays are then zeros.
struct Geometry1D {
  std::vector<double> dxi, lxi, dyj, lyj;
  bool quad = false;
  std::vector<double> q;
  // Rectilinear grid whose interior cells all have one dxi, lxi, dyj and lyj:
  // StageArgs::ug of those cells (uniform_cell_geometry)
  bool uniform = false;
  bool u_pow2 = false;
  double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
int build_geometry(int ni, int nj, int nh, const double *nx, const double *ny, Geometry1D &g);
int default_tile(const cato_config &c, int ni, int nj);
int validate_stage_args(const StageArgs &s);

// One solve: configuration, geometry and the device arrays it owns.
struct Solve {
  cato_config cfg{};
  int ni = 0, nj = 0, nh = 0, pitch = 0;
  int tile = 0, tiles_x = 0, tiles_y = 0;
  Geometry1D geo;
  StageLauncher launcher = nullptr;
  // device arrays (carved from one allocation by bind())
  double *U[3][4] = {};
  double *P[3][4] = {};      // primitive buffer sets; P[0] = primitives of U^n between steps
                             // (precision form: P[0] only, refreshed on demand, see prims_stale)
  double *dxi = nullptr, *lxi = nullptr, *dyj = nullptr, *lyj = nullptr;
  double *quad = nullptr;    // Q_COUNT per-cell geometry arrays (general quadrilateral grids only)
  double *block_min = nullptr;
  double *result = nullptr;
  int *block_err = nullptr;
  LoopState *loop = nullptr;
  double *src_E = nullptr;  // energy source field (pitched like U; owned by the caller of bind())
  double *pi_tab = nullptr; // pressure input table: n times, pressures, densities (owned by the caller)
  int pi_n = 0;
  // convergence diagnostics (owned by the caller of bind()): 8 arrays of block
  // sums, the 4 norms of the last step, and a ring of history rows
  double *conv_sums = nullptr, *conv_norms = nullptr, *conv_hist = nullptr;
  static constexpr int kHistCap = 64;
  // sub-domain of a split solve: sides that take their ghosts from the halo store (owned by the split)
  int halo_mask = 0;
  double *halo = nullptr;
  int cur = 0;             // which U[] holds U^n
  bool prims_valid = false; // P[0] holds the caller's primitives (not yet checked by a stage)
  bool prims_stale = false; // precision form: stages left P[0] behind U[cur] (they store no primitives)

  // allow_quad: accept a general quadrilateral grid (single solves; the batched,
  // served and split forms take rectilinear grids only)
  int init(const cato_config &c, const cato_geom &g, bool allow_quad = false);
  size_t quad_elems() const { return geo.quad ? (size_t)(ni + 2 * nh) * (nj + 2 * nh) : 0; }
  size_t field_elems() const;
  size_t device_bytes() const;
  void bind(char *base);
  int upload_geometry(hipStream_t s);
  FieldGeom fgeom() const;
  StageArgs stage_args(int mode, double a, double b, double c, int in_buf, int u0_buf, int out_buf, int p_in,
                       int p_out, bool last, const LoopState *lp, double dt_value, bool check_input) const;
  // keep_un: the last stage writes a free state buffer and primitive set instead
  // of U^n and P[0] (a step that may be retried from U^n, subcycle_check)
  std::vector<StageArgs> step_stages(const LoopState *lp, double dt_value, bool keep_un = false) const;
  ConvOut conv_out() const { return ConvOut{conv_sums, conv_norms, conv_hist, kHistCap}; }
};

// Host <-> device copy of one interior field with host layout (ld, halo).
// A tightly packed host field (ld == ni, halo == 0) moves as ONE contiguous
// copy, repacked on the device through `staging` (a free device field of the
// same solve) when the device pitch differs; other layouts use a 2D copy.
int derive_primitives(hipStream_t st, const Solve &s);
// +x pressure input of batched and served solves: the pass before stage m of a step, and the checks of a
// table and of the constant drive
PressurePass stage_pressure_pass(const Solve &s, const StageArgs &a, int m);
int check_pressure_table(const double *time, const double *pressure, const double *density, int n);
int check_pressure_constants(const Solve &s);
// Column ni+1 of a primitive set for a +x outflow side (no-op otherwise).
int fill_outflow_column(hipStream_t st, const Solve &s, double *const *P);
// Precision form: P[0] from U[cur] when stages have run since it was last set.
int refresh_prims(hipStream_t st, Solve &s);
int field_up(hipStream_t st, const Solve &s, double *dev, const double *host, int ld, int halo, double *staging);
int field_down(hipStream_t st, const Solve &s, double *host, const double *dev, int ld, int halo, double *staging);

}  // namespace cato
