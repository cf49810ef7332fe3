"""ctypes front end of the C oracle (oracle/cato_oracle.c).

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` use it as the checker.  The product path never
imports it.

Fields are numpy float64 arrays of shape (ni + 2*nh, nj + 2*nh) in Fortran
order, so element [i + nh - 1, j + nh - 1] is the reference's data(i, j)
(src/lib/field/field_2d.f90:477-483).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CATO_ORACLE_OMP=1 selects the OpenMP build (rows in parallel, same results),
# used only to time the CPU baseline on all host cores.
_OMP = os.environ.get("CATO_ORACLE_OMP") == "1"
LIB_PATH = os.path.join(HERE, "liboracle_omp.so" if _OMP else "liboracle.so")

FLUX_IDS = {
    "AUSMPW+": 0, "SLAU": 1, "SLAU2": 2, "SD-SLAU": 3, "SD-SLAU2": 4,
    "Roe": 5, "S-Roe": 5, "SP-Roe": 5, "AUSM+": 6, "AUSM+-u": 6, "HLLC": 7, "M-AUSMPW+": 8,
}
LIMITER_IDS = {"minmod": 0, "superbee": 1, "van_leer": 2, "vanleer": 2, "MLP3": 3, "MLP5": 4}
BC_IDS = {"periodic": 0, "zero_gradient": 1, "symmetry": 2, "mirror": 2, "reflection": 2, "outflow": 3, "outlet": 3,
          "pressure_input": 4}
RK_IDS = {
    "ssp_rk2": 0, "ssp_rk3": 1, "ssp_rk33": 1,
    "rk(1,3)": 2, "rk(1,4)": 3, "rk(1,5)": 4, "rk(2,3)": 5, "rk(2,4)": 6, "rk(2,5)": 7,
}


class _Config(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nj", C.c_int), ("nh", C.c_int),
        ("flux", C.c_int), ("limiter", C.c_int), ("integrator", C.c_int),
        ("bc_plus_x", C.c_int), ("bc_minus_x", C.c_int), ("bc_plus_y", C.c_int), ("bc_minus_y", C.c_int),
        ("apply_low_mach_fix", C.c_int),
        ("gamma", C.c_double), ("cfl", C.c_double), ("bc_pressure", C.c_double),
        ("bc_density", C.c_double), ("pi_n", C.c_int), ("pi_t", C.POINTER(C.c_double)),
        ("pi_p", C.POINTER(C.c_double)), ("pi_rho", C.POINTER(C.c_double)), ("bc_time", C.c_double),
        ("allow_subcycle", C.c_int), ("subcycle_tolerance", C.c_double),
    ]


class _Geometry(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nj", C.c_int), ("nh", C.c_int),
        ("volume", C.POINTER(C.c_double)), ("dx", C.POINTER(C.c_double)), ("dy", C.POINTER(C.c_double)),
        ("edge_lengths", C.POINTER(C.c_double)), ("edge_norm", C.POINTER(C.c_double)),
    ]


class _Fluid(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("rho", "rho_u", "rho_v", "rho_E", "u", "v", "p", "cs")] + \
        [("src_E", C.POINTER(C.c_double)), ("conv", C.c_double * 4), ("time", C.c_double), ("t0", C.c_double),
         ("t_solver", C.c_double), ("last_dt", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with the reference-faithful flags (no FMA contraction)."""
    src = os.path.join(HERE, "cato_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE, os.path.basename(LIB_PATH)])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.or_geometry_from_nodes.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.POINTER(_Geometry)]
        L.or_geometry_from_nodes.restype = C.c_int
        L.or_apply_bcs.argtypes = [C.POINTER(_Config), dp, dp, dp, dp]
        L.or_apply_bc_single.argtypes = [C.POINTER(_Config), C.c_int, C.c_int, dp, dp, dp, dp]
        L.or_interpolate_edges.argtypes = [C.POINTER(_Config), dp, dp, dp]
        L.or_edge_fluxes.argtypes = [C.POINTER(_Config), C.POINTER(_Geometry), dp, dp, dp, dp, dp, dp, dp]
        L.or_flux_split_edges.argtypes = [C.POINTER(_Config), C.POINTER(_Geometry), dp, dp, dp, dp, dp, dp]
        L.or_solve.argtypes = [C.POINTER(_Config), C.POINTER(_Geometry), dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_prim_to_cons.argtypes = [C.POINTER(_Config), dp, dp, dp, dp, dp, dp, dp]
        L.or_derived_quantities.argtypes = [C.POINTER(_Config), dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_fluid_init_from_prims.argtypes = [C.POINTER(_Config), C.POINTER(_Fluid)]
        L.or_get_timestep.argtypes = [C.POINTER(_Config), C.POINTER(_Geometry), C.POINTER(_Fluid)]
        L.or_get_timestep.restype = C.c_double
        L.or_integrate.argtypes = [C.POINTER(_Config), C.POINTER(_Geometry), C.POINTER(_Fluid), C.c_double]
        L.or_integrate.restype = C.c_int
        L.or_run.argtypes = [C.POINTER(_Config), C.POINTER(_Geometry), C.POINTER(_Fluid), C.c_double, C.c_long,
                             C.c_double, dp, C.POINTER(C.c_int)]
        L.or_run.restype = C.c_long
        L.or_interp_linear.argtypes = [C.c_int, dp, dp, C.c_double]
        L.or_interp_linear.restype = C.c_double
        _lib = L
    return _lib


def use_openmp() -> bool:
    """Switch this process to the OpenMP build (same results, rows in parallel) for large grids."""
    global _lib, LIB_PATH
    omp = os.path.join(HERE, "liboracle_omp.so")
    if LIB_PATH == omp:
        return True
    if not os.path.exists(omp):
        return False
    LIB_PATH, _lib = omp, None
    lib()
    return True


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Config:
    """Scheme configuration, with the reference deck's names (mod_input.f90:34-155)."""

    def __init__(self, ni, nj, flux="AUSMPW+", limiter="minmod", integrator="ssp_rk2",
                 bcs=("periodic", "periodic", "periodic", "periodic"), gamma=1.4, cfl=0.1,
                 apply_low_mach_fix=True, n_ghost=None, bc_pressure=0.0, bc_density=0.0, pressure_table=None,
                 bc_time=0.0, allow_subcycle=False, subcycle_tolerance=1e-3):
        self.ni, self.nj = int(ni), int(nj)
        self.bc_pressure = float(bc_pressure)
        # pressure_input: constant bc_pressure / bc_density, or pressure_table = (time, pressure, density)
        self.bc_density = float(bc_density)
        self.bc_time = float(bc_time)
        self.allow_subcycle, self.subcycle_tolerance = bool(allow_subcycle), float(subcycle_tolerance)
        self.pressure_table = None
        if pressure_table is not None:
            self.pressure_table = tuple(np.ascontiguousarray(a, dtype=np.float64) for a in pressure_table)
        self.flux, self.limiter, self.integrator = flux, limiter, integrator
        self.bcs = tuple(bcs)  # (+x, -x, +y, -y)
        self.gamma, self.cfl = float(gamma), float(cfl)
        self.apply_low_mach_fix = bool(apply_low_mach_fix)
        self.nh = n_ghost if n_ghost is not None else (3 if limiter == "MLP5" else 2)

    @property
    def shape(self):
        return (self.ni + 2 * self.nh, self.nj + 2 * self.nh)

    def c(self) -> _Config:
        tab = self.pressure_table
        dp = C.POINTER(C.c_double)
        return _Config(self.ni, self.nj, self.nh, FLUX_IDS[self.flux], LIMITER_IDS[self.limiter],
                       RK_IDS[self.integrator], BC_IDS[self.bcs[0]], BC_IDS[self.bcs[1]], BC_IDS[self.bcs[2]],
                       BC_IDS[self.bcs[3]], int(self.apply_low_mach_fix), self.gamma, self.cfl, self.bc_pressure,
                       self.bc_density, len(tab[0]) if tab else 0, _p(tab[0]) if tab else dp(),
                       _p(tab[1]) if tab else dp(), _p(tab[2]) if tab else dp(), self.bc_time,
                       int(self.allow_subcycle), self.subcycle_tolerance)

    def field(self):
        return np.zeros(self.shape, dtype=np.float64, order="F")


class Geometry:
    """Per-cell geometry of the halo'd block (grid_block_2d.f90:352-391)."""

    def __init__(self, cfg: Config, node_x: np.ndarray, node_y: np.ndarray):
        ni, nj, nh = cfg.ni, cfg.nj, cfg.nh
        nx, ny = cfg.shape
        assert node_x.shape == (nx + 1, ny + 1)
        self.node_x = np.asfortranarray(node_x, dtype=np.float64)
        self.node_y = np.asfortranarray(node_y, dtype=np.float64)
        self.volume = np.zeros((nx, ny), order="F")
        self.dx = np.zeros((nx, ny), order="F")
        self.dy = np.zeros((nx, ny), order="F")
        self.edge_lengths = np.zeros((4, nx, ny), order="F")
        self.edge_norm = np.zeros((2, 4, nx, ny), order="F")
        self._g = _Geometry(ni, nj, nh, _p(self.volume), _p(self.dx), _p(self.dy), _p(self.edge_lengths),
                            _p(self.edge_norm))
        rc = lib().or_geometry_from_nodes(ni, nj, nh, _p(self.node_x), _p(self.node_y), C.byref(self._g))
        if rc != 0:
            raise ValueError("negative or zero cell volume in the grid")


def uniform_nodes(ni, nj, nh, xrange=(0.0, 1.0), yrange=(0.0, 1.0)):
    """Node coordinates exactly as pycato.make_uniform_grid builds them
    (pycato/generate_initial_grids.py:43-112): np.linspace including ghost nodes."""
    dx = float(xrange[1] - xrange[0]) / float(ni)
    dy = float(yrange[1] - yrange[0]) / float(nj)
    x = np.linspace(xrange[0] - nh * dx, xrange[1] + nh * dx, (ni + 1) + 2 * nh, dtype=np.float64)
    y = np.linspace(yrange[0] - nh * dy, yrange[1] + nh * dy, (nj + 1) + 2 * nh, dtype=np.float64)
    y2, x2 = np.meshgrid(y, x)
    return np.asfortranarray(x2), np.asfortranarray(y2)


def interp_linear(t, f, x: float) -> float:
    """linear_interp_1d%evaluate (finterp's interp_1d) at x over knots t."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().or_interp_linear(len(t), _p(t), _p(f), float(x))


def apply_bcs(cfg: Config, rho, u, v, p):
    c = cfg.c()
    lib().or_apply_bcs(C.byref(c), _p(rho), _p(u), _p(v), _p(p))


LOCATIONS = {"+x": 0, "-x": 1, "+y": 2, "-y": 3}


def apply_bc_single(cfg: Config, bc_type: str, location: str, rho, u, v, p):
    """One boundary_condition_t%apply call (boundary_conditions.f90:60-70)."""
    c = cfg.c()
    lib().or_apply_bc_single(C.byref(c), BC_IDS[bc_type], LOCATIONS[location], _p(rho), _p(u), _p(v), _p(p))


def interpolate_edges(cfg: Config, q):
    ni, nj = cfg.ni, cfg.nj
    ie = np.zeros((2, ni + 1, nj + 1), order="F")
    je = np.zeros((2, ni + 1, nj + 1), order="F")
    c = cfg.c()
    lib().or_interpolate_edges(C.byref(c), _p(q), _p(ie), _p(je))
    return ie, je


def edge_fluxes(cfg: Config, geom: Geometry, rho, u, v, p, edges):
    """edges: list of 8 arrays (rho_i, rho_j, u_i, u_j, v_i, v_j, p_i, p_j), each (2, ni+1, nj+1);
    for M-AUSMPW+ 16: the same 8 followed by the 8 superbee reconstructions."""
    ni, nj = cfg.ni, cfg.nj
    packed = np.concatenate([np.asfortranarray(e).ravel(order="F") for e in edges])
    iflux = np.zeros((4, ni + 1, nj), order="F")
    jflux = np.zeros((4, ni, nj + 1), order="F")
    c = cfg.c()
    lib().or_edge_fluxes(C.byref(c), C.byref(geom._g), _p(rho), _p(u), _p(v), _p(p), _p(packed), _p(iflux),
                         _p(jflux))
    return iflux, jflux


def flux_split_edges(cfg: Config, geom: Geometry, iflux, jflux):
    d = [cfg.field() for _ in range(4)]
    c = cfg.c()
    lib().or_flux_split_edges(C.byref(c), C.byref(geom._g), _p(np.asfortranarray(iflux)),
                              _p(np.asfortranarray(jflux)), *[_p(x) for x in d])
    return d


def solve(cfg: Config, geom: Geometry, rho, u, v, p):
    """flux_solver_t%solve: BCs applied in place to (rho, u, v, p); returns d(rho, rho u, rho v, rho E)/dt."""
    d = [cfg.field() for _ in range(4)]
    c = cfg.c()
    lib().or_solve(C.byref(c), C.byref(geom._g), _p(rho), _p(u), _p(v), _p(p), *[_p(x) for x in d])
    return d


class Fluid:
    """fluid_t state of the oracle: conserved + primitive halo'd fields."""

    NAMES = ("rho", "rho_u", "rho_v", "rho_E", "u", "v", "p", "cs")

    def __init__(self, cfg: Config, geom: Geometry, rho, u, v, p):
        self.cfg, self.geom = cfg, geom
        for n in self.NAMES:
            setattr(self, n, cfg.field())
        nh, ni, nj = cfg.nh, cfg.ni, cfg.nj
        sl = (slice(nh, nh + ni), slice(nh, nh + nj))
        for name, src in (("rho", rho), ("u", u), ("v", v), ("p", p)):
            getattr(self, name)[sl] = np.asarray(src, dtype=np.float64)[sl] if src.shape == cfg.shape else src
        self._f = _Fluid(*[_p(getattr(self, n)) for n in self.NAMES], None)
        self._src = None
        c = cfg.c()
        lib().or_fluid_init_from_prims(C.byref(c), C.byref(self._f))

    def set_energy_source(self, src):
        """Halo'd field added to d(rho E)/dt in every stage (None removes it)."""
        if src is None:
            self._src = None
            self._f.src_E = None
        else:
            self._src = np.asfortranarray(src, dtype=np.float64)
            assert self._src.shape == self.cfg.shape
            self._f.src_E = _p(self._src)

    def set_time(self, t: float):
        """fluid_t time and the time its flux solver starts from (a fluid initialized at time t)."""
        self._f.time = float(t)
        self._f.t0 = float(t)
        self._f.t_solver = float(t)

    @property
    def time(self) -> float:
        return float(self._f.time)

    @property
    def last_dt(self) -> float:
        """The dt the last integrate() used (smaller than the one given after subcycling)."""
        return float(self._f.last_dt)

    def convergence(self):
        """L2-norm convergence [rho, rho u, rho v, rho E] of the last integrate()."""
        return list(self._f.conv)

    def get_timestep(self) -> float:
        c = self.cfg.c()
        return lib().or_get_timestep(C.byref(c), C.byref(self.geom._g), C.byref(self._f))

    def integrate(self, dt: float) -> int:
        c = self.cfg.c()
        return lib().or_integrate(C.byref(c), C.byref(self.geom._g), C.byref(self._f), float(dt))

    def run(self, max_time: float, max_iterations: int = 2 ** 31 - 1, contour_interval_dt: float = 0.1):
        c = self.cfg.c()
        t = C.c_double(0.0)
        err = C.c_int(0)
        n = lib().or_run(C.byref(c), C.byref(self.geom._g), C.byref(self._f), float(max_time), int(max_iterations),
                         float(contour_interval_dt), C.byref(t), C.byref(err))
        return int(n), float(t.value), int(err.value)

    def interior(self, name):
        nh, ni, nj = self.cfg.nh, self.cfg.ni, self.cfg.nj
        return getattr(self, name)[nh:nh + ni, nh:nh + nj].copy(order="F")
