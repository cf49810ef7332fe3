"""One solve split into px x py sub-domains on one GPU (cato_split_*), halos exchanged after every RK stage
by the pack / unpack kernels, the dt the minimum over the sub-domains: bit for bit the unsplit solve, for
periodic wraps between sub-domains, physical sides meeting exchanged ones at corners, uneven blocks and
three ghost layers (MLP5)."""
import sys

import numpy as np
import pytest

from fvleg_2d_amd import SchemeConfig, Solver
from fvleg_2d_amd.problems import kelvin_helmholtz, riemann2d, uniform_nodes
from fvleg_2d_amd.solver import SplitSolver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def smooth_state(ni, nj, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(ni) + 0.5) / ni
    y = (np.arange(nj) + 0.5) / nj
    X, Y = np.meshgrid(x, y, indexing="ij")
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (X + rng.uniform())) * np.cos(2 * np.pi * (Y + rng.uniform()))
    rho[(X - 0.4) ** 2 + (Y - 0.6) ** 2 < 0.04] = 2.5
    u = 0.3 * np.cos(2 * np.pi * (Y + rng.uniform()))
    v = 0.25 * np.sin(2 * np.pi * (X + 2 * Y + rng.uniform()))
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (X - Y + rng.uniform()))
    p[(X - 0.4) ** 2 + (Y - 0.6) ** 2 < 0.04] = 3.0
    return [np.asfortranarray(a) for a in (rho, u, v, p)]


def compare(cfg, nx, ny, prims, px, py, steps):
    with Solver(cfg, nx, ny) as s, SplitSolver(cfg, nx, ny, px, py) as sp:
        s.set_primitive(*prims)
        sp.set_primitive(*prims)
        for k in range(steps):
            dt, dts = s.compute_dt(), sp.compute_dt()
            assert dt == dts, (k, dt, dts)
            assert s.step(dt) == 0 and sp.step(dts) == 0
        a, b = s.conserved(), sp.conserved()
        ms, n = sp.exchange_time()
    assert n == steps * (3 if cfg.integrator == "ssp_rk3" else 2)  # one exchange per stage
    for v in range(4):
        assert np.array_equal(a[v], b[v]), v
    say("  exchanges", n, "GPU ms per exchange", ms / max(n, 1))


@pytest.mark.parametrize("px, py", [(4, 2), (2, 3), (1, 4)])
def test_split_periodic_kh_is_bit_identical(px, py):
    say("split KH", px, py)
    pb = kelvin_helmholtz(n=128, nj=96)
    compare(pb.cfg, pb.node_x, pb.node_y, [pb.rho, pb.u, pb.v, pb.p], px, py, 8)


@pytest.mark.parametrize("flux, lim, bcs, px, py", [
    ("HLLC", "MLP5", ("zero_gradient",) * 4, 3, 2),
    ("SLAU", "van_leer", ("symmetry", "symmetry", "periodic", "periodic"), 2, 3),
    ("AUSMPW+", "minmod", ("zero_gradient", "symmetry", "symmetry", "zero_gradient"), 3, 4),
    ("Roe", "superbee", ("periodic", "periodic", "symmetry", "zero_gradient"), 4, 3),
])
def test_split_with_physical_sides_is_bit_identical(flux, lim, bcs, px, py):
    """Uneven blocks (100 x 60 over 3 or 4 columns / rows), corners where an exchanged side meets a
    physical one, and the SSP-RK2 stage cycle."""
    say("split", flux, lim, bcs, px, py)
    ni, nj = 100, 60
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk2", bcs=bcs, cfl=0.4)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost)
    compare(cfg, nx, ny, smooth_state(ni, nj, 3), px, py, 6)


def test_split_riemann_rk3_mlp5_is_bit_identical():
    pb = riemann2d(n=96, nj=72, seed=5)
    compare(pb.cfg, pb.node_x, pb.node_y, [pb.rho, pb.u, pb.v, pb.p], 3, 2, 5)


def test_split_rejects_what_it_does_not_cover():
    pb = kelvin_helmholtz(n=64, nj=64, precision="fp32_flux")
    with pytest.raises(Exception):
        SplitSolver(pb.cfg, pb.node_x, pb.node_y, 2, 2)
    pb = kelvin_helmholtz(n=64, nj=64)
    with pytest.raises(Exception):
        SplitSolver(pb.cfg, pb.node_x, pb.node_y, 64, 1)  # blocks narrower than the halo
