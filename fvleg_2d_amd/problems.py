"""Seeded synthetic initial conditions of the benchmark workloads.

Grids are built exactly as pycato's ``make_uniform_grid`` builds them
(pycato/generate_initial_grids.py:43-112: np.linspace over the domain plus
ghost nodes, meshgrid with i fastest), so node coordinates and hence the
per-cell geometry match what the reference reads from an input deck.
Random draws use numpy's PCG64 generator (``np.random.default_rng(seed)``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .solver import SchemeConfig


def uniform_nodes(ni: int, nj: int, nh: int, xrange=(0.0, 1.0), yrange=(0.0, 1.0)):
    dx = float(xrange[1] - xrange[0]) / float(ni)
    dy = float(yrange[1] - yrange[0]) / float(nj)
    x = np.linspace(xrange[0] - nh * dx, xrange[1] + nh * dx, (ni + 1) + 2 * nh, dtype=np.float64)
    y = np.linspace(yrange[0] - nh * dy, yrange[1] + nh * dy, (nj + 1) + 2 * nh, dtype=np.float64)
    y2, x2 = np.meshgrid(y, x)
    return np.asfortranarray(x2), np.asfortranarray(y2)


def wavy_nodes(ni: int, nj: int, nh: int, amp=0.04, xrange=(0.0, 1.0), yrange=(0.0, 1.0)):
    """A curvilinear grid: the uniform nodes displaced by amp * (domain length) * sin(2 pi x) sin(2 pi y)
    (x) and sin(2 pi x + 1) sin(2 pi y) (y), periodic in both directions, cells skewed by up to
    atan(2 pi amp) (14 degrees at amp = 0.04) and every volume positive for amp < 0.11."""
    x, y = uniform_nodes(ni, nj, nh, xrange, yrange)
    lx, ly = float(xrange[1] - xrange[0]), float(yrange[1] - yrange[0])
    xn, yn = (x - xrange[0]) / lx, (y - yrange[0]) / ly
    s = np.sin(2 * np.pi * yn)
    xw = x + amp * lx * np.sin(2 * np.pi * xn) * s
    yw = y + amp * ly * np.sin(2 * np.pi * xn + 1.0) * s
    return np.asfortranarray(xw), np.asfortranarray(yw)


def cell_centers(node_x, node_y, nh):
    """Interior cell centres as pycato computes them (x_2d[:-1,:-1] + dx/2)."""
    xc = 0.5 * (node_x[:-1, :-1] + node_x[1:, :-1])
    yc = 0.5 * (node_y[:-1, :-1] + node_y[:-1, 1:])
    return xc[nh:-nh, nh:-nh], yc[nh:-nh, nh:-nh]


@dataclass
class Problem:
    name: str
    cfg: SchemeConfig
    node_x: np.ndarray
    node_y: np.ndarray
    rho: np.ndarray
    u: np.ndarray
    v: np.ndarray
    p: np.ndarray
    max_time: float = 0.0
    max_iterations: int = 0
    contour_interval_dt: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def ni(self):
        return self.rho.shape[0]

    @property
    def nj(self):
        return self.rho.shape[1]


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def sod(ni=400, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk2", cfl=0.4, max_time=0.2,
        precision="fp64") -> Problem:
    """Sod shock tube on [0,1]x[0,0.01] (the sod_1d deck's states, tests/integrated/sod_1d/generate_ic.py)."""
    cfg = SchemeConfig(flux=flux, limiter=limiter, integrator=integrator,
                       bcs=("zero_gradient", "zero_gradient", "periodic", "periodic"), gamma=1.4, cfl=cfl,
                       precision=precision)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost, (0.0, 1.0), (0.0, 0.01))
    xc, _ = cell_centers(nx, ny, cfg.n_ghost)
    left = xc < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    return Problem("sod", cfg, nx, ny, _f(rho), _f(z), _f(z.copy()), _f(p), max_time=max_time,
                   contour_interval_dt=max_time)


def riemann_states(seed: int):
    rng = np.random.default_rng(seed)
    st = []
    for _ in range(4):
        st.append((rng.uniform(0.1, 2.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(0.1, 1.5)))
    return st


def riemann2d(n=1024, seed=1, flux="HLLC", limiter="MLP5", integrator="ssp_rk3", cfl=0.4, max_time=0.25,
              precision="fp64", nj=None) -> Problem:
    """Four-quadrant Riemann problem on [0,1]^2 with seeded quadrant states."""
    nj = n if nj is None else nj
    cfg = SchemeConfig(flux=flux, limiter=limiter, integrator=integrator,
                       bcs=("zero_gradient",) * 4, gamma=1.4, cfl=cfl, precision=precision)
    nx, ny = uniform_nodes(n, nj, cfg.n_ghost)
    xc, yc = cell_centers(nx, ny, cfg.n_ghost)
    st = riemann_states(seed)
    q = np.where(xc >= 0.5, np.where(yc >= 0.5, 0, 3), np.where(yc >= 0.5, 1, 2))
    arr = np.array(st)
    rho, u, v, p = (arr[q, k] for k in range(4))
    return Problem(f"riemann2d_{n}x{nj}", cfg, nx, ny, _f(rho), _f(u), _f(v), _f(p), max_time=max_time,
                   contour_interval_dt=max_time, meta={"seed": seed, "states": st})


def kelvin_helmholtz(n=2048, seed=7, flux="SLAU", limiter="van_leer", integrator="ssp_rk3", cfl=0.5, steps=2000,
                     precision="fp64", nj=None) -> Problem:
    """Periodic shear layer: rho=2, u=0.5 for |y-0.5|<0.25, rho=1, u=-0.5 elsewhere, p=2.5;
    v = 0.01 * sum of 8 random-phase sine modes in x plus 1e-3 white noise.  nj (default n) rows of
    the same square cells: the domain is [0,1] x [0, nj/n]."""
    nj = n if nj is None else nj
    cfg = SchemeConfig(flux=flux, limiter=limiter, integrator=integrator, bcs=("periodic",) * 4, gamma=1.4,
                       cfl=cfl, precision=precision)
    nx, ny = uniform_nodes(n, nj, cfg.n_ghost, (0.0, 1.0), (0.0, nj / n))
    xc, yc = cell_centers(nx, ny, cfg.n_ghost)
    rng = np.random.default_rng(seed)
    inner = np.abs(yc - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, 0.5, -0.5)
    phases = rng.uniform(0.0, 2.0 * math.pi, 8)
    v = np.zeros_like(xc)
    for m in range(8):
        v += np.sin(2.0 * math.pi * (m + 1) * xc + phases[m])
    v = 0.01 * v + 1e-3 * rng.uniform(-1.0, 1.0, xc.shape)
    p = np.full_like(xc, 2.5)
    return Problem(f"kelvin_helmholtz_{n}x{nj}", cfg, nx, ny, _f(rho), _f(u), _f(v), _f(p), max_time=1e300,
                   max_iterations=steps, contour_interval_dt=1e300, meta={"seed": seed})


BATCH_FLUXES = ("AUSM+", "AUSMPW+", "HLLC", "Roe")
BATCH_LIMITERS = ("minmod", "MLP5")


def batch_riemann(n_solves=256, seed=2024, side_range=(64, 512), max_time=0.2, precision="fp64",
                  integrator="ssp_rk3"):
    """Independent four-quadrant Riemann solves: square side log-uniform in side_range, quadrant
    states as in riemann2d, flux from {AUSM+, AUSMPW+, HLLC, Roe}, reconstruction from
    {MUSCL-minmod, MLP5}."""
    rng = np.random.default_rng(seed)
    lo, hi = side_range
    out = []
    for k in range(n_solves):
        side = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        flux = BATCH_FLUXES[int(rng.integers(len(BATCH_FLUXES)))]
        lim = BATCH_LIMITERS[int(rng.integers(len(BATCH_LIMITERS)))]
        pb = riemann2d(n=side, seed=int(rng.integers(1 << 30)), flux=flux, limiter=lim, integrator=integrator,
                       max_time=max_time, precision=precision)
        pb.name = f"batch{k}_{side}_{flux}_{lim}"
        out.append(pb)
    return out
