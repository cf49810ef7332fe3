"""General quadrilateral (curvilinear) grids on the MI355X: per-cell volumes, edge lengths, min_dx / min_dy
and edge normals (quad_cell.f90:69-229) read by the fused stage kernel and the flux operator, bit for bit
the oracle's arithmetic (whose geometry and normal terms tests/test_oracle_quad.py pins), on a wavy grid
whose cells are skewed by up to 20 degrees: edge fluxes of all nine solvers, residuals and SSP-RK3 steps
over the limiters, boundary conditions, the device-resident loop, the FP32-flux form's error, and the
refusal of the forms that take rectilinear grids only."""
import dataclasses

import numpy as np
import pytest

from fvleg_2d_amd import SchemeConfig, Solver
from fvleg_2d_amd.problems import riemann2d, wavy_nodes
from test_gpu_parity import FLUXES, oracle_cfg, pad, rel_linf, run_both, run_steps, say, shock_state

pytestmark = pytest.mark.gpu

AMP = 0.055


@pytest.mark.parametrize("flux", FLUXES)
def test_quad_edge_flux_parity(oracle, flux):
    say("quad edge_flux", flux)
    ni, nj = 19, 15
    cfg = SchemeConfig(flux=flux, limiter="minmod")
    nh = cfg.n_ghost
    nx, ny = wavy_nodes(ni, nj, nh, AMP)
    rho, u, v, p = (pad(a, nh) for a in shock_state(ni, nj, 3))
    oc = oracle_cfg(oracle, cfg, ni, nj)
    oracle.apply_bcs(oc, rho, u, v, p)
    edges = []
    for q in (rho, u, v, p):
        edges += list(oracle.interpolate_edges(oc, q))
    if flux == "M-AUSMPW+":
        osb = oracle_cfg(oracle, SchemeConfig(flux=flux, limiter="superbee"), ni, nj)
        for q in (rho, u, v, p):
            edges += list(oracle.interpolate_edges(osb, q))
    og = oracle.Geometry(oc, nx, ny)
    oi, oj = oracle.edge_fluxes(oc, og, rho, u, v, p, edges)
    with Solver(cfg, nx, ny) as s:
        gi, gj = s.edge_fluxes(edges, p)
    assert np.array_equal(gi, oi), np.abs(gi - oi).max()
    assert np.array_equal(gj, oj), np.abs(gj - oj).max()


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("lim", ["minmod", "van_leer", "MLP5"])
def test_quad_residual_parity(oracle, flux, lim):
    say("quad residual", flux, lim)
    ni, nj = 37, 29  # partial 16 x 16 tiles on both sides
    cfg = SchemeConfig(flux=flux, limiter=lim, bcs=("zero_gradient", "symmetry", "periodic", "periodic"))
    nh = cfg.n_ghost
    nx, ny = wavy_nodes(ni, nj, nh, AMP, (0.0, 1.0), (0.0, 0.8))
    prims = [pad(a, nh) for a in shock_state(ni, nj, 4)]
    oc = oracle_cfg(oracle, cfg, ni, nj)
    og = oracle.Geometry(oc, nx, ny)
    o_in = [a.copy(order="F") for a in prims]
    od = oracle.solve(oc, og, *o_in)
    with Solver(cfg, nx, ny) as s:
        g_in = [a.copy(order="F") for a in prims]
        gd = s.residual(*g_in)
    for a, b in zip(g_in, o_in):  # the boundary fill, in place
        assert np.array_equal(a, b)
    for k in range(4):
        assert np.array_equal(gd[k], od[k]), (k, rel_linf(gd[k], od[k]))


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("lim", ["superbee", "MLP3"])
def test_quad_step_parity_ssp_rk3(oracle, flux, lim):
    """Six SSP-RK3 steps: CFL dt from min_dx / min_dy equal every step, state and primitives bit for bit."""
    say("quad step", flux, lim)
    ni, nj = 45, 38
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk3", bcs=("zero_gradient",) * 4, cfl=0.4)
    nx, ny = wavy_nodes(ni, nj, cfg.n_ghost, AMP)
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, shock_state(ni, nj, 5), 6)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    for k in range(5):
        assert np.array_equal(gp[k], op[k]), ("prim", k)


@pytest.mark.parametrize("flux, lim, bcs", [
    ("SLAU2", "MLP5", ("symmetry", "symmetry", "periodic", "periodic")),
    ("AUSMPW+", "van_leer", ("outflow", "zero_gradient", "symmetry", "symmetry")),
    ("HLLC", "minmod", ("periodic", "periodic", "zero_gradient", "symmetry")),
    ("Roe", "MLP3", ("zero_gradient", "outflow", "periodic", "periodic")),
])
def test_quad_step_parity_boundaries(oracle, flux, lim, bcs):
    say("quad step bcs", flux, lim, bcs)
    ni, nj = 33, 26
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk2", bcs=bcs, cfl=0.4, bc_pressure=0.6)
    nx, ny = wavy_nodes(ni, nj, cfg.n_ghost, AMP, (0.0, 1.0), (0.0, 0.7))
    rho, u, v, p = shock_state(ni, nj, 8)
    if "outflow" in bcs:
        u = np.asfortranarray(u + 0.4)
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, [rho, u, v, p], 8)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))


def quad_riemann(n, precision="fp64"):
    pb = riemann2d(n=n, seed=3, precision=precision)
    nx, ny = wavy_nodes(n, n, pb.cfg.n_ghost, AMP)
    return dataclasses.replace(pb, node_x=nx, node_y=ny, max_time=0.05, contour_interval_dt=0.05)


def test_quad_device_loop_matches_oracle(oracle):
    """The device-resident loop (graph replays, dt from the last stage's CFL partials) on a 72 x 72 wavy
    grid, HLLC + MLP5, to t = 0.05: iterations, time and state equal to the oracle's run."""
    pb = quad_riemann(72)
    (n_g, t_g, e_g, g), (n_o, t_o, e_o, o) = run_both(oracle, pb)
    say("quad loop iterations", n_g, n_o, "time", t_g, t_o)
    assert e_g == e_o == 0 and n_g == n_o and t_g == t_o
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))


def test_quad_precision_form_error():
    """FP32 reconstruction and flux (normals rounded to FP32) against the FP64 path on a 256 x 256 wavy
    grid to t = 0.05: relative L1 <= 1e-5 and L-infinity <= 1e-2 per conserved variable."""
    out = {}
    for prec in ("fp64", "fp32_flux"):
        pb = quad_riemann(256, prec)
        with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            n, t, e = s.run(pb.max_time, 0, pb.max_time)
            assert e == 0 and t == pb.max_time
            out[prec] = s.conserved()
    for k in range(4):
        r, g = out["fp64"][k], out["fp32_flux"][k]
        scale = np.abs(r).max()
        l1, linf = np.mean(np.abs(g - r)) / scale, np.max(np.abs(g - r)) / scale
        say("  var", k, "rel L1", l1, "Linf", linf)
        assert l1 <= 1e-5 and linf <= 1e-2


def test_quad_grid_is_refused_by_the_rectilinear_forms():
    from fvleg_2d_amd.batch import Batch
    from fvleg_2d_amd.solver import SplitSolver

    pb = quad_riemann(48)
    with pytest.raises(Exception, match="quadrilateral"):
        Batch([pb])
    with pytest.raises(Exception, match="quadrilateral"):
        SplitSolver(pb.cfg, pb.node_x, pb.node_y, 2, 2)
