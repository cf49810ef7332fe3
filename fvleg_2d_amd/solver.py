"""Python handle on one device-resident CATO solve (a ``cato_ctx``).

Fields cross the boundary as numpy float64 arrays in Fortran order, shaped
either (ni, nj) (interior only) or (ni + 2*nh, nj + 2*nh) (the reference's
halo'd ``field_2d_t%data`` layout, src/lib/field/field_2d.f90:477-483).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class SchemeConfig:
    """The [scheme]/[time]/[physics]/[boundary_conditions] deck entries the hot path needs
    (src/lib/io/mod_input.f90:34-155)."""

    def __init__(self, flux="AUSMPW+", limiter="minmod", integrator="ssp_rk2",
                 bcs=("periodic", "periodic", "periodic", "periodic"), gamma=1.4, cfl=0.1,
                 apply_low_mach_fix=True, precision="fp64", tile=-1, bc_pressure=0.0, bc_density=0.0):
        self.flux, self.limiter, self.integrator = flux, limiter, integrator
        # outflow_ambient_pressure, or the pressure_input's constant_bc_pressure_value
        self.bc_pressure = float(bc_pressure)
        self.bc_density = float(bc_density)  # pressure_input: bc_density
        self.bcs = tuple(bcs)  # (+x, -x, +y, -y)
        self.gamma, self.cfl = float(gamma), float(cfl)
        self.apply_low_mach_fix = bool(apply_low_mach_fix)
        self.precision = precision
        self.tile = int(tile)

    @property
    def n_ghost(self) -> int:
        return 3 if self.limiter == "MLP5" else 2

    def to_c(self) -> N.CatoConfig:
        try:
            return N.CatoConfig(N.FLUX[self.flux], N.LIMITER[self.limiter], N.INTEGRATOR[self.integrator],
                                N.PRECISION[self.precision], N.BC[self.bcs[0]], N.BC[self.bcs[1]],
                                N.BC[self.bcs[2]], N.BC[self.bcs[3]], self.n_ghost, int(self.apply_low_mach_fix),
                                self.gamma, self.cfl, self.tile, self.bc_pressure, self.bc_density)
        except KeyError as e:
            raise ValueError(f"unknown scheme option {e}") from None


class Solver:
    """One solve resident in HBM: ``cato_ctx_create`` ... ``cato_ctx_destroy``."""

    def __init__(self, cfg: SchemeConfig, node_x: np.ndarray, node_y: np.ndarray):
        self.lib = N.load()
        self.cfg = cfg
        self.node_x, self.node_y = _f64f(node_x), _f64f(node_y)
        nh = cfg.n_ghost
        self.ni = self.node_x.shape[0] - 1 - 2 * nh
        self.nj = self.node_x.shape[1] - 1 - 2 * nh
        self.nh = nh
        self._ccfg = cfg.to_c()
        g = N.CatoGeom(self.ni, self.nj, _p(self.node_x), _p(self.node_y))
        h = C.c_void_p()
        N.check(self.lib.cato_ctx_create(C.byref(self._ccfg), C.byref(g), C.byref(h)), "cato_ctx_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.cato_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- layout helpers ------------------------------------------------------
    def _ld_halo(self, a: np.ndarray):
        if a.shape == (self.ni, self.nj):
            return self.ni, 0
        if a.shape == (self.ni + 2 * self.nh, self.nj + 2 * self.nh):
            return self.ni + 2 * self.nh, self.nh
        raise ValueError(f"field shape {a.shape} is neither interior nor halo'd")

    def halo_shape(self):
        return (self.ni + 2 * self.nh, self.nj + 2 * self.nh)

    # ---- state ---------------------------------------------------------------
    def set_primitive(self, rho, u, v, p):
        arrs = [_f64f(x) for x in (rho, u, v, p)]
        ld, halo = self._ld_halo(arrs[0])
        N.check(self.lib.cato_set_primitive(self.h, *[_p(a) for a in arrs], ld, halo), "cato_set_primitive")

    def set_conserved(self, rho, rho_u, rho_v, rho_E):
        arrs = [_f64f(x) for x in (rho, rho_u, rho_v, rho_E)]
        ld, halo = self._ld_halo(arrs[0])
        N.check(self.lib.cato_set_conserved(self.h, *[_p(a) for a in arrs], ld, halo), "cato_set_conserved")

    def set_energy_source(self, q_dot):
        """Energy source added to d(rho E)/dt in every stage of the following steps
        (fluid_t%time_derivative, mod_fluid.f90:500-509); None clears it."""
        if q_dot is None:
            N.check(self.lib.cato_set_energy_source(self.h, None, self.ni, 0), "cato_set_energy_source")
            return
        q = _f64f(q_dot)
        ld, halo = self._ld_halo(q)
        N.check(self.lib.cato_set_energy_source(self.h, _p(q), ld, halo), "cato_set_energy_source")

    def set_pressure_input(self, table):
        """Time-dependent pressure input (pressure_input_bc_t%read_pressure_input): table = (time, pressure,
        density) rows, already scaled; None returns to the constant bc_pressure / bc_density."""
        if table is None:
            N.check(self.lib.cato_set_pressure_input(self.h, None, None, None, 0), "cato_set_pressure_input")
            return
        t, p, r = (np.ascontiguousarray(a, dtype=np.float64) for a in table)
        if not (len(t) == len(p) == len(r)):
            raise ValueError("time, pressure and density need the same length")
        N.check(self.lib.cato_set_pressure_input(self.h, _p(t), _p(p), _p(r), len(t)), "cato_set_pressure_input")

    def set_subcycle(self, on: bool, tolerance: float = 1e-3):
        """[time] allow_subcycle / subcycle_error_tolerance: step() retries with dt / 10 while a convergence
        norm exceeds the tolerance (SSP-RK2/3, host-driven steps)."""
        N.check(self.lib.cato_set_subcycle(self.h, int(on), float(tolerance)), "cato_set_subcycle")

    def set_loop_dt(self, initial_dt: float = 0.0, startup: bool = False, startup_period: float = 0.0,
                    startup_dt: float = 0.0, start_iteration: int = 0, exact_landing: bool = False):
        """main.f90:140-158 dt choices of run(): startup_dt while the driver time <= startup_period (startup),
        else initial_dt on the driver's iteration 0 (start_iteration + the run's steps), else CFL/constant.
        exact_landing (not in the reference): steps are shortened so the state lands on the contour times and
        max_time."""
        N.check(self.lib.cato_set_loop_dt(self.h, float(initial_dt), int(bool(startup)), float(startup_period),
                                          float(startup_dt), int(start_iteration), int(bool(exact_landing))),
                "cato_set_loop_dt")

    def last_dt(self) -> float:
        """The dt the last step() used."""
        dt = C.c_double()
        N.check(self.lib.cato_last_dt(self.h, C.byref(dt)), "cato_last_dt")
        return dt.value

    def set_time(self, t: float):
        """Fluid time (and the time the solve starts from): fluid_t%set_time on an initialized fluid."""
        N.check(self.lib.cato_set_time(self.h, float(t)), "cato_set_time")

    def conserved(self):
        out = [np.zeros((self.ni, self.nj), order="F") for _ in range(4)]
        N.check(self.lib.cato_get_conserved(self.h, *[_p(a) for a in out], self.ni, 0), "cato_get_conserved")
        return out

    def primitive(self):
        out = [np.zeros((self.ni, self.nj), order="F") for _ in range(5)]
        N.check(self.lib.cato_get_primitive(self.h, *[_p(a) for a in out], self.ni, 0), "cato_get_primitive")
        return out  # rho, u, v, p, cs

    # ---- time integration ----------------------------------------------------
    def compute_dt(self) -> float:
        dt = C.c_double()
        N.check(self.lib.cato_compute_dt(self.h, C.byref(dt)), "cato_compute_dt")
        return dt.value

    def step(self, dt: float) -> int:
        err = C.c_int()
        N.check(self.lib.cato_step(self.h, float(dt), C.byref(err)), "cato_step")
        return err.value

    def run(self, max_time: float, max_iterations: int = 0, contour_interval_dt: float = 0.0,
            constant_dt: float = 0.0):
        it, t, err = C.c_longlong(), C.c_double(), C.c_int()
        N.check(self.lib.cato_run(self.h, float(max_time), int(max_iterations), float(contour_interval_dt),
                                  float(constant_dt), C.byref(it), C.byref(t), C.byref(err)), "cato_run")
        return it.value, t.value, err.value

    def advance(self, n: int):
        N.check(self.lib.cato_advance(self.h, int(n)), "cato_advance")

    def loop_status(self):
        """(steps taken, driver time, error code) of the last run / advance, after waiting for it."""
        it, t, err = C.c_longlong(), C.c_double(), C.c_int()
        N.check(self.lib.cato_loop_status(self.h, C.byref(it), C.byref(t), C.byref(err)), "cato_loop_status")
        return it.value, t.value, err.value

    def synchronize(self):
        N.check(self.lib.cato_synchronize(self.h), "cato_synchronize")

    # ---- step bookkeeping (calculate_convergence / write_residual_history) ------
    def set_diagnostics(self, on: bool):
        N.check(self.lib.cato_set_diagnostics(self.h, int(on)), "cato_set_diagnostics")

    def convergence(self):
        """L2-norm convergence [rho, rho u, rho v, rho E] of the last step."""
        out = (C.c_double * 4)()
        N.check(self.lib.cato_get_convergence(self.h, out), "cato_get_convergence")
        return list(out)

    def history(self) -> np.ndarray:
        """Collected residual_hist rows (iteration, time, dt, 4 norms), oldest first; drains them."""
        n = C.c_longlong()
        N.check(self.lib.cato_get_history(self.h, None, 0, C.byref(n)), "cato_get_history")
        rows = np.zeros((n.value, 7))
        if n.value:
            N.check(self.lib.cato_get_history(self.h, _p(rows), n.value, C.byref(n)), "cato_get_history")
        return rows[: n.value]

    def set_device_graph(self, on: bool):
        """run() / advance() replay a captured 32-step hipGraph (on, the default) or launch every kernel."""
        N.check(self.lib.cato_set_device_graph(self.h, int(on)), "cato_set_device_graph")

    def set_timing(self, on: bool):
        N.check(self.lib.cato_set_timing(self.h, int(on)), "cato_set_timing")

    def timing(self):
        a, b, c, d = C.c_double(), C.c_longlong(), C.c_double(), C.c_longlong()
        N.check(self.lib.cato_get_timing(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "cato_get_timing")
        return {"stage_ms": a.value, "stage_launches": b.value, "other_ms": c.value, "other_launches": d.value}

    # ---- operator-level entry points ----------------------------------------
    def apply_bc(self, rho, u, v, p):
        """In place on halo'd Fortran arrays (boundary_condition_t%apply of every side)."""
        for a in (rho, u, v, p):
            assert a.shape == self.halo_shape() and a.flags.f_contiguous and a.dtype == np.float64
        N.check(self.lib.cato_apply_bc(self.h, _p(rho), _p(u), _p(v), _p(p), self.halo_shape()[0]), "cato_apply_bc")

    def interpolate_edges(self, q):
        q = _f64f(q)
        assert q.shape == self.halo_shape()
        ie = np.zeros((2, self.ni + 1, self.nj + 1), order="F")
        je = np.zeros((2, self.ni + 1, self.nj + 1), order="F")
        N.check(self.lib.cato_interpolate_edges(self.h, _p(q), q.shape[0], _p(ie), _p(je)), "cato_interpolate_edges")
        return ie, je

    def edge_fluxes(self, edges, p):
        packed = np.concatenate([_f64f(e).ravel(order="F") for e in edges])
        p = _f64f(p)
        fi = np.zeros((4, self.ni + 1, self.nj), order="F")
        fj = np.zeros((4, self.ni, self.nj + 1), order="F")
        N.check(self.lib.cato_edge_fluxes(self.h, _p(packed), _p(p), p.shape[0], _p(fi), _p(fj)), "cato_edge_fluxes")
        return fi, fj

    def residual(self, rho, u, v, p):
        """flux_solver_t%solve: BCs applied in place to the halo'd primitives; returns d/dt of the
        conserved variables as halo'd arrays with zero halos."""
        for a in (rho, u, v, p):
            assert a.shape == self.halo_shape() and a.flags.f_contiguous and a.dtype == np.float64
        d = [np.zeros(self.halo_shape(), order="F") for _ in range(4)]
        N.check(self.lib.cato_residual(self.h, _p(rho), _p(u), _p(v), _p(p), *[_p(x) for x in d],
                                       self.halo_shape()[0]), "cato_residual")
        return d


class SplitSolver:
    """One solve split into px x py sub-domains on one GPU with halos exchanged after every RK stage
    (``cato_split_*``): the decomposition a run over several GPUs uses, bit-identical to ``Solver``."""

    def __init__(self, cfg: SchemeConfig, node_x: np.ndarray, node_y: np.ndarray, px: int, py: int):
        self.lib = N.load()
        self.cfg = cfg
        self.node_x, self.node_y = _f64f(node_x), _f64f(node_y)
        nh = cfg.n_ghost
        self.ni = self.node_x.shape[0] - 1 - 2 * nh
        self.nj = self.node_x.shape[1] - 1 - 2 * nh
        self._ccfg = cfg.to_c()
        g = N.CatoGeom(self.ni, self.nj, _p(self.node_x), _p(self.node_y))
        h = C.c_void_p()
        N.check(self.lib.cato_split_create(C.byref(self._ccfg), C.byref(g), int(px), int(py), C.byref(h)),
                "cato_split_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.cato_split_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_primitive(self, rho, u, v, p):
        arrs = [_f64f(a) for a in (rho, u, v, p)]
        N.check(self.lib.cato_split_set_primitive(self.h, *[_p(a) for a in arrs], self.ni, 0),
                "cato_split_set_primitive")

    def conserved(self):
        out = [np.zeros((self.ni, self.nj), order="F") for _ in range(4)]
        N.check(self.lib.cato_split_get_conserved(self.h, *[_p(a) for a in out], self.ni, 0),
                "cato_split_get_conserved")
        return out

    def compute_dt(self) -> float:
        dt = C.c_double()
        N.check(self.lib.cato_split_compute_dt(self.h, C.byref(dt)), "cato_split_compute_dt")
        return dt.value

    def step(self, dt: float) -> int:
        err = C.c_int()
        N.check(self.lib.cato_split_step(self.h, float(dt), C.byref(err)), "cato_split_step")
        return err.value

    def exchange_time(self):
        """(GPU ms of all halo exchanges so far, number of exchanges)."""
        ms, n = C.c_double(), C.c_longlong()
        N.check(self.lib.cato_split_exchange_time(self.h, C.byref(ms), C.byref(n)), "cato_split_exchange_time")
        return ms.value, n.value
