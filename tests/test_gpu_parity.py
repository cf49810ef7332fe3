"""Parity of the HIP path against the oracle (run on an MI355X).

Every comparison goes through the C ABI of libcato_hip.so.  FP64 results are
expected to be bit-identical to the oracle; the asserted bar is the
north-star tolerance, relative L-infinity <= 1e-12 per conserved variable.
"""
import sys

import numpy as np
import pytest

from fvleg_2d_amd import SchemeConfig, Solver
from error_cases import ERROR_CASES, error_case_fields, error_case_oracle
from error_cases import NI as EC_NI, NJ as EC_NJ
from fvleg_2d_amd.problems import kelvin_helmholtz, riemann2d, sod, uniform_nodes

pytestmark = pytest.mark.gpu

FLUXES = ["AUSMPW+", "SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2", "Roe", "AUSM+", "HLLC", "M-AUSMPW+"]
LIMITERS = ["minmod", "superbee", "van_leer", "MLP3", "MLP5"]
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_linf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def oracle_cfg(O, cfg: SchemeConfig, ni, nj):
    return O.Config(ni, nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                    gamma=cfg.gamma, cfl=cfg.cfl, apply_low_mach_fix=cfg.apply_low_mach_fix,
                    bc_pressure=cfg.bc_pressure)


def pad(a, nh):
    return np.asfortranarray(np.pad(np.asarray(a, dtype=np.float64), nh))


def shock_state(ni, nj, seed):
    """Smooth random field with an embedded oblique jump (exercises limiters and sensors)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(ni) + 0.5) / ni
    y = (np.arange(nj) + 0.5) / nj
    X, Y = np.meshgrid(x, y, indexing="ij")
    jump = (X + 0.3 * Y) > 0.55
    rho = np.where(jump, 0.4, 1.2) + 0.05 * np.sin(2 * np.pi * (X + rng.uniform()))
    u = 0.3 * np.cos(2 * np.pi * (Y + rng.uniform())) + np.where(jump, 0.2, -0.1)
    v = 0.25 * np.sin(2 * np.pi * (X + 2 * Y + rng.uniform()))
    p = np.where(jump, 0.3, 1.0) + 0.02 * np.cos(2 * np.pi * (X - Y + rng.uniform()))
    return [np.asfortranarray(a) for a in (rho, u, v, p)]


BC_SETS = [
    ("zero_gradient",) * 4,
    ("periodic",) * 4,
    ("zero_gradient", "zero_gradient", "periodic", "periodic"),
    ("periodic", "periodic", "zero_gradient", "zero_gradient"),
    ("symmetry",) * 4,
    ("symmetry", "zero_gradient", "zero_gradient", "symmetry"),
    ("zero_gradient", "symmetry", "periodic", "periodic"),
    ("outflow", "zero_gradient", "symmetry", "symmetry"),
    ("outflow", "symmetry", "periodic", "periodic"),
    ("outflow",) * 4,
    ("symmetry", "outflow", "outflow", "zero_gradient"),
]


@pytest.mark.parametrize("bcs", BC_SETS, ids=lambda b: "-".join(x[:3] for x in b))
@pytest.mark.parametrize("lim", ["minmod", "MLP5"])
def test_apply_bc_parity(oracle, bcs, lim):
    say("apply_bc", bcs, lim)
    ni, nj = 13, 9
    cfg = SchemeConfig(limiter=lim, bcs=bcs, bc_pressure=0.9)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh)
    rng = np.random.default_rng(1)
    # rho, p > 0 and |u|, |v| up to 2: both branches of the +x outflow fill (subsonic and supersonic rows)
    lo_hi = [(0.5, 2.0), (-2.0, 2.0), (-2.0, 2.0), (0.2, 2.0)]
    fields = [np.asfortranarray(rng.uniform(a, b, (ni + 2 * nh, nj + 2 * nh))) for a, b in lo_hi]
    with Solver(cfg, nx, ny) as s:
        g = [f.copy(order="F") for f in fields]
        s.apply_bc(*g)
    oc = oracle_cfg(oracle, cfg, ni, nj)
    o = [f.copy(order="F") for f in fields]
    oracle.apply_bcs(oc, *o)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("lim", LIMITERS)
def test_interpolate_edges_parity(oracle, lim):
    say("interpolate", lim)
    ni, nj = 21, 17
    cfg = SchemeConfig(limiter=lim)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh)
    rng = np.random.default_rng(2)
    q = np.asfortranarray(np.cumsum(rng.uniform(-1, 1, (ni + 2 * nh, nj + 2 * nh)), axis=0))
    q[5:9, :] += 3.0
    with Solver(cfg, nx, ny) as s:
        gi, gj = s.interpolate_edges(q)
    oi, oj = oracle.interpolate_edges(oracle_cfg(oracle, cfg, ni, nj), q)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gj, oj)


@pytest.mark.parametrize("flux", FLUXES)
def test_edge_flux_parity(oracle, flux):
    say("edge_flux", flux)
    ni, nj = 19, 15
    cfg = SchemeConfig(flux=flux, limiter="minmod")
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh)
    rho, u, v, p = (pad(a, nh) for a in shock_state(ni, nj, 3))
    oc = oracle_cfg(oracle, cfg, ni, nj)
    oracle.apply_bcs(oc, rho, u, v, p)
    edges = []
    for q in (rho, u, v, p):
        edges += list(oracle.interpolate_edges(oc, q))
    if flux == "M-AUSMPW+":  # plus the superbee reconstruction it blends in
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
@pytest.mark.parametrize("lim", LIMITERS)
def test_residual_parity(oracle, flux, lim):
    say("residual", flux, lim)
    ni, nj = 37, 29  # not multiples of any tile: partial tiles on both edges
    cfg = SchemeConfig(flux=flux, limiter=lim, bcs=("zero_gradient", "symmetry", "periodic", "periodic"))
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh, (0.0, 1.0), (0.0, 0.8))
    prims = [pad(a, nh) for a in shock_state(ni, nj, 4)]
    oc = oracle_cfg(oracle, cfg, ni, nj)
    og = oracle.Geometry(oc, nx, ny)
    o_in = [a.copy(order="F") for a in prims]
    od = oracle.solve(oc, og, *o_in)
    with Solver(cfg, nx, ny) as s:
        g_in = [a.copy(order="F") for a in prims]
        gd = s.residual(*g_in)
    for a, b in zip(g_in, o_in):
        assert np.array_equal(a, b)
    for k in range(4):
        assert np.array_equal(gd[k], od[k]), (k, rel_linf(gd[k], od[k]))


def run_steps(oracle, cfg, nx, ny, prims, nsteps):
    ni, nj = prims[0].shape
    nh = cfg.n_ghost
    oc = oracle_cfg(oracle, cfg, ni, nj)
    og = oracle.Geometry(oc, nx, ny)
    f = oracle.Fluid(oc, og, *[pad(a, nh) for a in prims])
    with Solver(cfg, nx, ny) as s:
        s.set_primitive(*prims)
        for k in range(nsteps):
            dt_g = s.compute_dt()
            dt_o = f.get_timestep()
            assert dt_g == dt_o, (k, dt_g, dt_o)
            eg = s.step(dt_g)
            eo = f.integrate(dt_o)
            assert eg == eo == 0
        g = s.conserved()
        gp = s.primitive()
    o = [f.interior(n) for n in ("rho", "rho_u", "rho_v", "rho_E")]
    op = [f.interior(n) for n in ("rho", "u", "v", "p", "cs")]
    return g, o, gp, op


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("lim", LIMITERS)
def test_step_parity_ssp_rk3(oracle, flux, lim):
    say("step rk3", flux, lim)
    ni, nj = 45, 38
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk3", bcs=("zero_gradient",) * 4, cfl=0.4)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost)
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, shock_state(ni, nj, 5), 6)
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    for k in range(5):
        assert np.array_equal(gp[k], op[k]), ("prim", k)


def rectilinear_nodes(xs, ys):
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return np.asfortranarray(X), np.asfortranarray(Y)


# (ni, nj, x nodes, y nodes, every interior cell alike): exact binary spacings take the epilogue's
# host-computed uniform geometry (power-of-two volume 2^-12; volume 0.5625 with equal edge lengths; dx != dy),
# a stretched grid the per-cell arrays
GEOM_GRIDS = {
    "pow2": (64, 32, lambda nh: (np.arange(-nh, 64 + nh + 1) / 64.0, np.arange(-nh, 32 + nh + 1) / 64.0), True),
    "square_0.75": (48, 36, lambda nh: (np.arange(-nh, 48 + nh + 1) * 0.75, np.arange(-nh, 36 + nh + 1) * 0.75), True),
    "dx_ne_dy": (40, 32, lambda nh: (np.arange(-nh, 40 + nh + 1) / 32.0, np.arange(-nh, 32 + nh + 1) * (3.0 / 128.0)),
                 True),
    "stretched": (40, 32, lambda nh: (np.cumsum(np.r_[0.0, 0.02 + 0.01 * np.sin(np.arange(40 + 2 * nh))]) - 0.05,
                                      np.cumsum(np.r_[0.0, 0.03 + 0.005 * np.cos(np.arange(32 + 2 * nh))]) - 0.07),
                  False),
}


@pytest.mark.parametrize("grid", list(GEOM_GRIDS))
@pytest.mark.parametrize("flux, lim", [("SLAU", "van_leer"), ("HLLC", "MLP5"), ("AUSMPW+", "minmod")])
def test_rectilinear_cell_geometry_step_parity(oracle, grid, flux, lim):
    """SSP-RK3 steps on rectilinear grids whose interior cells are all alike (the stage epilogue takes the
    cell geometry computed once on the host) and on a stretched grid (per-cell arrays): CFL dt, state and
    primitives bit for bit the oracle's."""
    ni, nj, nodes, uniform = GEOM_GRIDS[grid]
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk3", bcs=("zero_gradient",) * 4, cfl=0.4)
    xs, ys = nodes(cfg.n_ghost)
    nh = cfg.n_ghost
    dx, dy = np.diff(xs)[nh:nh + ni], np.diff(ys)[nh:nh + nj]
    assert (np.all(dx == dx[0]) and np.all(dy == dy[0])) == uniform
    nx, ny = rectilinear_nodes(xs, ys)
    say("rectilinear geometry", grid, flux, lim)
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, shock_state(ni, nj, 11), 6)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    for k in range(5):
        assert np.array_equal(gp[k], op[k]), ("prim", k)


def fast_ops_edge_states(ni, nj, seed):
    """States around the operand ranges of the FP64 FastOps edge arithmetic (fast_ops_cell: rho and p in
    [2^-B, 2^B], u and v zero or 2^-150 <= |u| <= 2^B; B = 60 for TVD2 with the SLAU family, 20 with MLP or
    HLLC).  'scaled': density and pressure of every cell times 1e20, so no tile qualifies; 'scaled_1e7':
    times 1e7, beyond 2^20 but within 2^60; 'mixed': a patch of velocities 1e-300 (whose quotients need the
    IEEE range scaling) in some tiles and negative-zero velocities (allowed) in others, next to tiles that
    qualify; 'wide': density over [2^-19, 2^19] at pressure 1 (sound speeds over 2^19), every tile within
    the B = 20 bound."""
    rho, u, v, p = shock_state(ni, nj, seed)
    scaled = [rho * 1e20, u.copy(order="F"), v.copy(order="F"), p * 1e20]
    scaled7 = [rho * 1e7, u.copy(order="F"), v.copy(order="F"), p * 1e7]
    u2, v2 = u.copy(order="F"), v.copy(order="F")
    u2[4:12, 3:9] = 1e-300
    v2[4:12, 3:9] = -1e-300
    u2[40:56, 20:30] = -0.0
    v2[40:56, 20:30] = -0.0
    mixed = [rho.copy(order="F"), u2, v2, p.copy(order="F")]
    X, Y = np.meshgrid((np.arange(ni) + 0.5) / ni, (np.arange(nj) + 0.5) / nj, indexing="ij")
    rw = np.exp2(19.0 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y))
    wide = [np.asfortranarray(rw), np.asfortranarray(0.05 * u), np.asfortranarray(0.05 * v),
            np.asfortranarray(np.ones_like(rw))]
    return {"scaled": scaled, "scaled_1e7": scaled7, "mixed": mixed, "wide": wide}


@pytest.mark.parametrize("flux", ["SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2", "HLLC", "AUSM+", "AUSMPW+", "M-AUSMPW+",
                                  "Roe"])
@pytest.mark.parametrize("lim", ["minmod", "van_leer", "MLP3", "MLP5"])
@pytest.mark.parametrize("case", ["scaled", "scaled_1e7", "mixed", "wide"])
def test_fast_ops_tiles_and_their_fallback(oracle, flux, lim, case):
    """The stage kernel takes its FastOps edge arithmetic (the reconstruction of every scheme, the flux of
    the SLAU family and HLLC) only on tiles whose cells keep it in range and the IEEE sequences elsewhere:
    both bit-identical to the oracle, with symmetry ghosts (mirrored velocities, negative zeros) at two
    sides."""
    say("fast ops", flux, lim, case)
    ni, nj = 64, 48
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk3",
                       bcs=("symmetry", "zero_gradient", "periodic", "periodic"), cfl=0.4)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost)
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, fast_ops_edge_states(ni, nj, 8)[case], 4)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    for k in range(5):
        assert np.array_equal(gp[k], op[k]), ("prim", k)


@pytest.mark.parametrize("flux, lim", [("AUSMPW+", "minmod"), ("SLAU", "MLP5"), ("M-AUSMPW+", "van_leer")])
def test_step_parity_outflow(oracle, flux, lim):
    """outlet_bc.f90 at +x (subsonic extrapolation to the ambient pressure) with symmetry in y, 20 steps."""
    say("step outflow", flux, lim)
    ni, nj = 47, 21
    cfg = SchemeConfig(flux=flux, limiter=lim, integrator="ssp_rk3",
                       bcs=("outflow", "zero_gradient", "symmetry", "symmetry"), cfl=0.4, bc_pressure=0.6)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost, (0.0, 1.0), (0.0, 0.5))
    rho, u, v, p = shock_state(ni, nj, 8)
    u = np.asfortranarray(u + 0.4)  # flow out through +x
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, [rho, u, v, p], 20)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    for k in range(5):
        assert np.array_equal(gp[k], op[k]), ("prim", k)


@pytest.mark.parametrize("integrator", ["ssp_rk2", "rk(1,3)", "rk(2,4)", "rk(2,5)"])
@pytest.mark.parametrize("flux", ["SLAU", "AUSMPW+"])
def test_step_parity_integrators(oracle, integrator, flux):
    say("step", integrator, flux)
    ni, nj = 33, 20
    cfg = SchemeConfig(flux=flux, limiter="MLP3", integrator=integrator, bcs=("periodic",) * 4, cfl=0.3)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost, (0.0, 1.0), (0.0, 0.6))
    g, o, _, _ = run_steps(oracle, cfg, nx, ny, shock_state(ni, nj, 6), 5)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(2,4)"])
def test_energy_source_parity(oracle, integrator):
    """time_derivative's energy source (mod_fluid.f90:500-509): a non-uniform q_dot held over several steps,
    then cleared; residual() (solver%solve) must ignore it."""
    say("energy source", integrator)
    ni, nj = 41, 27
    cfg = SchemeConfig(flux="SLAU2", limiter="MLP5", integrator=integrator,
                       bcs=("zero_gradient", "symmetry", "periodic", "periodic"), cfl=0.35)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh, (0.0, 1.0), (0.0, 0.7))
    prims = shock_state(ni, nj, 11)
    rng = np.random.default_rng(12)
    X, Y = np.meshgrid((np.arange(ni) + 0.5) / ni, (np.arange(nj) + 0.5) / nj, indexing="ij")
    q = np.asfortranarray(40.0 * np.exp(-((X - 0.4) ** 2 + (Y - 0.3) ** 2) / 0.02) + rng.uniform(0, 1, (ni, nj)))
    oc = oracle_cfg(oracle, cfg, ni, nj)
    og = oracle.Geometry(oc, nx, ny)
    f = oracle.Fluid(oc, og, *[pad(a, nh) for a in prims])
    f.set_energy_source(pad(q, nh))
    with Solver(cfg, nx, ny) as s:
        s.set_primitive(*prims)
        s.set_energy_source(pad(q, nh))  # halo'd layout; the interior is what counts
        for k in range(6):
            if k == 4:
                f.set_energy_source(None)
                s.set_energy_source(None)
            dt = s.compute_dt()
            assert dt == f.get_timestep(), k
            assert s.step(dt) == 0 and f.integrate(dt) == 0
        s.set_energy_source(q)  # interior layout
        g = s.conserved()
        gp = s.primitive()
        rp = [pad(a, nh) for a in gp[:4]]
        gd = s.residual(*[a.copy(order="F") for a in rp])
    o = [f.interior(n) for n in ("rho", "rho_u", "rho_v", "rho_E")]
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    od = oracle.solve(oc, og, *[a.copy(order="F") for a in rp])
    for k in range(4):
        assert np.array_equal(gd[k], od[k]), ("residual", k)


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(2,4)"])
def test_convergence_history_parity(oracle, integrator):
    """calculate_convergence + write_residual_history (mod_fluid.f90:1162-1223): per-step norms within 1e-12 of
    the oracle's NORM2 (a different summation order), rows from cato_step and from the device loop, and the
    state still bit-identical with the diagnostics on."""
    say("convergence", integrator)
    ni, nj = 53, 31
    cfg = SchemeConfig(flux="AUSMPW+", limiter="superbee", integrator=integrator,
                       bcs=("zero_gradient", "zero_gradient", "periodic", "periodic"), cfl=0.4)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh, (0.0, 1.0), (0.0, 0.6))
    prims = shock_state(ni, nj, 21)
    oc = oracle_cfg(oracle, cfg, ni, nj)
    og = oracle.Geometry(oc, nx, ny)

    def oracle_rows(n):
        f = oracle.Fluid(oc, og, *[pad(a, nh) for a in prims])
        rows, t = [], 0.0
        for k in range(n):
            dt = f.get_timestep()
            assert f.integrate(dt) == 0
            t = t + dt
            rows.append([k + 1, t, dt] + f.convergence())
        return np.array(rows), f

    nsteps = 40  # crosses one 32-step poll of the device loop
    want, f = oracle_rows(nsteps)
    with Solver(cfg, nx, ny) as s:
        s.set_primitive(*prims)
        s.set_diagnostics(True)
        for k in range(5):
            s.step(s.compute_dt())
            assert np.allclose(s.convergence(), want[k, 3:], rtol=1e-12, atol=0), k
        got = s.history()
        assert got.shape == (5, 7)
        assert np.array_equal(got[:, :3], want[:5, :3])
        assert np.allclose(got[:, 3:], want[:5, 3:], rtol=1e-12, atol=0)
        assert s.history().shape == (0, 7)  # drained

        s.set_primitive(*prims)
        it, _, err = s.run(1e30, nsteps)
        assert (it, err) == (nsteps, 0)
        got = s.history()
        g = s.conserved()
    assert got.shape == (nsteps, 7)
    assert np.array_equal(got[:, :3], want[:, :3])
    assert np.allclose(got[:, 3:], want[:, 3:], rtol=1e-12, atol=0)
    assert np.all(want[:, 3:] > 0)
    for k, name in enumerate(("rho", "rho_u", "rho_v", "rho_E")):
        assert np.array_equal(g[k], f.interior(name)), name


@pytest.mark.parametrize("precision", ["fp64", "fp32_flux"])
def test_device_loop_graph_matches_eager(oracle, precision):
    """cato_advance / cato_run replay a captured 32-step graph unless the graph is switched off or timing is on
    (then every kernel is launched on its own): identical states and history rows, also after the stage arguments change between
    runs (energy source on, diagnostics on: the graph is dropped and captured again).  FP64 also matches the
    oracle's run loop (main.f90:138-198) bit for bit."""
    say("device-loop graph", precision)
    ni, nj = 45, 33
    # outlet pressure near the interior's: the outlet's Mach switch (outlet_bc.f90:96-123) then never flips, and
    # a 1e-9 perturbation of the start stays at 1e-9 (at 0.8 it grows to O(1) within 20 steps, even in FP64)
    cfg = SchemeConfig(flux="SLAU", limiter="MLP5", integrator="ssp_rk3",
                       bcs=("outflow", "symmetry", "periodic", "periodic"), bc_pressure=0.5, cfl=0.4,
                       precision=precision)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh, (0.0, 1.0), (0.0, 0.8))
    prims = shock_state(ni, nj, 31)
    X, Y = np.meshgrid((np.arange(ni) + 0.5) / ni, (np.arange(nj) + 0.5) / nj, indexing="ij")
    q = np.asfortranarray(30.0 * np.exp(-((X - 0.5) ** 2 + (Y - 0.4) ** 2) / 0.03))

    def drive(mode):
        with Solver(cfg, nx, ny) as s:
            s.set_timing(mode == "timed")
            s.set_device_graph(mode != "eager")
            s.set_primitive(*prims)
            s.advance(70)  # one eager step after set_primitive, two 32-step replays, five eager steps
            first = s.conserved()
            s.set_energy_source(q)
            s.advance(40)
            s.set_diagnostics(True)
            it, t, err = s.run(1e30, 75)
            return first, s.conserved(), s.history(), (it, t, err), s.timing()["stage_launches"]

    g1, g2, gh, gr, g_launches = drive("graph")
    for mode in ("eager", "timed"):
        e1, e2, eh, er, e_launches = drive(mode)
        # cato_run launches whole 32-step polls (96 steps for 75; the steps after the last one are no-ops)
        assert g_launches == 0 and e_launches == (3 * (70 + 40 + 96) if mode == "timed" else 0)
        assert gr == er and gr[0] == 75 and gr[2] == 0
        assert np.array_equal(gh, eh) and gh.shape == (75, 7)
        for k in range(4):
            assert np.array_equal(g1[k], e1[k]), (mode, "advance", k)
            assert np.array_equal(g2[k], e2[k]), (mode, "after source + diagnostics", k)
    if precision == "fp64":
        oc = oracle_cfg(oracle, cfg, ni, nj)
        f = oracle.Fluid(oc, oracle.Geometry(oc, nx, ny), *[pad(a, nh) for a in prims])
        assert f.run(1e300, 70, 1e300)[0] == 70
        for k, name in enumerate(("rho", "rho_u", "rho_v", "rho_E")):
            assert np.array_equal(g1[k], f.interior(name)), name


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(1,3)"])
def test_subcycle_parity(oracle, integrator):
    """[time] allow_subcycle (subcycle_check, mod_fluid.f90:856-887): cato_step retries with dt / 10 while a
    convergence norm exceeds the tolerance, from the same state, and keeps the last attempt's buffers; steps and
    the dt each used match the oracle bit for bit (rk(1,3) has no retry loop).  The device loop refuses it."""
    say("subcycle", integrator)
    from fvleg_2d_amd.problems import sod
    from fvleg_2d_amd._native import CatoError

    pb = sod(ni=100, nj=4, flux="AUSM+", limiter="minmod", integrator=integrator, cfl=0.4)
    cfg = pb.cfg
    nh = cfg.n_ghost
    tol = 3e-3
    oc = oracle_cfg(oracle, cfg, pb.ni, pb.nj)
    oc.allow_subcycle, oc.subcycle_tolerance = True, tol
    f = oracle.Fluid(oc, oracle.Geometry(oc, pb.node_x, pb.node_y), *[pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    retried = 0
    with Solver(cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        s.set_subcycle(True, tol)
        s.set_diagnostics(True)
        for k in range(12):
            dt = s.compute_dt()
            assert dt == f.get_timestep(), k
            assert s.step(dt) == 0 and f.integrate(dt) == 0
            assert s.last_dt() == f.last_dt, k
            retried += s.last_dt() < dt
            assert np.allclose(s.convergence(), f.convergence(), rtol=1e-12, atol=0), k
        g = s.conserved()
        h = s.history()
        if integrator != "rk(1,3)":
            with pytest.raises(CatoError, match="host-driven"):
                s.run(1.0, 5)
    for k, name in enumerate(("rho", "rho_u", "rho_v", "rho_E")):
        assert np.array_equal(g[k], f.interior(name)), name
    assert retried > 0 if integrator != "rk(1,3)" else retried == 0
    assert h.shape == (12, 7) and h[-1, 1] == f.time


@pytest.mark.parametrize("case", ["kh2048", "riemann1024"])
def test_full_size_workloads_match_oracle(oracle, case):
    """The bench workloads at their full sizes: Kelvin-Helmholtz 2048^2 (SLAU, van Leer, SSP-RK3) for 3 steps
    and the four-quadrant Riemann problem 1024^2 (HLLC, MLP5, SSP-RK3) for 5 steps, bit-exact (OpenMP oracle)."""
    say("full size", case)
    oracle.use_openmp()
    pb = kelvin_helmholtz(n=2048) if case == "kh2048" else riemann2d(n=1024, seed=1, flux="HLLC", limiter="MLP5")
    g, o, gp, op = run_steps(oracle, pb.cfg, pb.node_x, pb.node_y, [pb.rho, pb.u, pb.v, pb.p],
                             3 if case == "kh2048" else 5)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))
    for k in range(5):
        assert np.array_equal(gp[k], op[k]), ("prim", k)


def run_both(oracle, pb, max_iterations=0):
    nh = pb.cfg.n_ghost
    oc = oracle_cfg(oracle, pb.cfg, pb.ni, pb.nj)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    mi = max_iterations if max_iterations else 2 ** 31 - 1
    n_o, t_o, e_o = f.run(pb.max_time, mi, pb.contour_interval_dt)
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        n_g, t_g, e_g = s.run(pb.max_time, max_iterations, pb.contour_interval_dt)
        g = s.conserved()
    o = [f.interior(n) for n in ("rho", "rho_u", "rho_v", "rho_E")]
    return (n_g, t_g, e_g, g), (n_o, t_o, e_o, o)


def test_sod_config1_full_run(oracle):
    """Sod 400x4, AUSM+, MUSCL-minmod, SSP-RK2, CFL 0.4, to t = 0.2 (first benchmark configuration)."""
    pb = sod()
    say("sod run", pb.ni, pb.nj)
    (n_g, t_g, e_g, g), (n_o, t_o, e_o, o) = run_both(oracle, pb)
    say("iterations", n_g, n_o, "time", t_g, t_o)
    assert e_g == e_o == 0
    assert n_g == n_o and t_g == t_o
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL
        assert np.array_equal(g[k], o[k])


def test_exact_landing_option(oracle):
    """Outside the reference (which clips only the driver time, main.f90:178-194): with exact_landing the
    device loop shortens the steps that would carry the state past the next contour time or max_time, so
    the state lands on them.  The oracle steps with the same clipped dt; states, iterations and the final
    time must agree bit for bit, and every contour time must be a step's end time."""
    pb = sod()
    nh = pb.cfg.n_ghost
    T = pb.max_time
    c = T / 4
    oc = oracle_cfg(oracle, pb.cfg, pb.ni, pb.nj)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    eps = np.finfo(float).eps
    master = time = 0.0
    nxt = 0.0 + c
    it, ends, targets = 0, [], [nxt]

    def landing(m):
        out = nxt if nxt - master > 0.0 else nxt + c
        if out - master > 0.0 and m > out - master:
            m = out - master
        if T - master > 0.0 and m > T - master:
            m = T - master
        return m

    while time < T:
        d = landing(f.get_timestep())
        assert f.integrate(d) == 0
        master = master + d
        new = master
        if abs(time - nxt) < eps:
            nxt = nxt + c
            targets.append(nxt)
        if new > nxt:
            new = time + abs(nxt - time)
        time = new
        it += 1
        ends.append(master)
    assert time == master == T
    assert all(t in ends for t in targets if t <= T)
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        s.set_diagnostics(True)
        s.set_loop_dt(exact_landing=True)
        n_g, t_g, e_g = s.run(T, 0, c)
        hist = s.history()
        g = s.conserved()
    say("exact landing", n_g, it, t_g)
    assert e_g == 0 and n_g == it and t_g == T
    assert np.array_equal(hist[:, 1], ends)
    for k, name in enumerate(("rho", "rho_u", "rho_v", "rho_E")):
        assert np.array_equal(g[k], f.interior(name)), name


def test_riemann2d_reduced_full_run(oracle):
    """Four-quadrant Riemann problem (HLLC, MLP5, SSP-RK3, CFL 0.4) to t = 0.25 at 96^2."""
    pb = riemann2d(n=96, seed=1)
    say("riemann2d run", pb.ni)
    (n_g, t_g, e_g, g), (n_o, t_o, e_o, o) = run_both(oracle, pb)
    say("iterations", n_g, n_o)
    assert e_g == e_o == 0 and n_g == n_o and t_g == t_o
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL


def test_kelvin_helmholtz_reduced(oracle):
    """KH shear layer (SLAU, van Leer, SSP-RK3, CFL 0.5, periodic) at 128^2 for 60 steps."""
    pb = kelvin_helmholtz(n=128, steps=60)
    say("kh run", pb.ni)
    (n_g, t_g, e_g, g), (n_o, t_o, e_o, o) = run_both(oracle, pb, max_iterations=60)
    assert n_g == n_o == 60 and t_g == t_o
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL


@pytest.mark.parametrize("lim", ["minmod", "MLP5"])
@pytest.mark.parametrize("bcs", BC_SETS[:7] + BC_SETS[8:9], ids=lambda b: "-".join(x[:3] for x in b))
def test_minimal_grid_parity(oracle, bcs, lim):
    """The smallest grid the reference accepts, ni = nj = n_ghost_layers (mod_input.f90:320-338): every halo
    cell maps across the whole interior (periodic wrap, symmetry mirror of the full width), 4 SSP-RK3 steps."""
    say("minimal grid", bcs, lim)
    cfg = SchemeConfig(flux="AUSMPW+", limiter=lim, integrator="ssp_rk3", bcs=bcs, cfl=0.3, bc_pressure=0.7)
    n = cfg.n_ghost
    nx, ny = uniform_nodes(n, n, n, (0.0, 0.1), (0.0, 0.1))
    rng = np.random.default_rng(31)
    prims = [np.asfortranarray(rng.uniform(lo, hi, (n, n))) for lo, hi in ((0.8, 1.2), (-0.2, 0.2),
                                                                          (-0.2, 0.2), (0.8, 1.2))]
    g, o, gp, op = run_steps(oracle, cfg, nx, ny, prims, 4)
    for k in range(4):
        assert np.array_equal(g[k], o[k]), (k, rel_linf(g[k], o[k]))


def test_maximum_size_grid_conserves_totals():
    """8192^2 cells (67 M, ~13 GB of device state) on a periodic grid: 3 SSP-RK3 steps keep the mass and
    momentum totals to round-off and the energy total to round-off (no source), the state finite."""
    say("maximum size")
    pb = kelvin_helmholtz(n=8192, steps=3)
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        u0 = s.conserved()
        it, _, err = s.run(1e300, 3, 1e300)
        u1 = s.conserved()
    assert (it, err) == (3, 0)
    for k in range(4):
        assert np.all(np.isfinite(u1[k]))
        tot0, tot1 = u0[k].sum(dtype=np.float64), u1[k].sum(dtype=np.float64)
        mag = np.abs(u0[k]).sum()
        say("8192^2 total", k, tot0, tot1)
        assert abs(tot1 - tot0) <= 1e-12 * mag


@pytest.mark.parametrize("name", sorted(ERROR_CASES))
@pytest.mark.parametrize("loop", ["step", "run"])
def test_error_codes_match_the_oracle(oracle, name, loop):
    """fluid_t%integrate's error code (error.f90:6-9) for a NaN planted in the input, a NaN produced by
    a stage, negative density, negative pressure and two at once, from cato_step and from the device
    loop (cato_run, one step of constant dt): the exact code of the oracle's or_integrate, and for the
    cases without NaN the same (finite, partly negative) state bit for bit."""
    c = ERROR_CASES[name]
    say("error code", name, loop)
    cfg = SchemeConfig(flux="SLAU", limiter="minmod", integrator=c["integrator"], bcs=("zero_gradient",) * 4)
    nx, ny = uniform_nodes(EC_NI, EC_NJ, cfg.n_ghost)
    code_o, f = error_case_oracle(oracle, name)
    with Solver(cfg, nx, ny) as s:
        s.set_primitive(*error_case_fields(name))
        if loop == "step":
            code = s.step(c["dt"])
        else:
            it, _, code = s.run(1e300, 1, 1e300, c["dt"])
            assert it == 1
        got = s.conserved()
    assert code == code_o == c["code"], (code, code_o, c["code"])
    if code != 3:
        for k, n in enumerate(("rho", "rho_u", "rho_v", "rho_E")):
            assert np.array_equal(got[k], f.interior(n)), n


def test_invalid_configuration_is_rejected():
    cfg = SchemeConfig(flux="SLAU", limiter="MLP5")
    cfg.tile = 99
    nx, ny = uniform_nodes(8, 8, 3)
    with pytest.raises(Exception):
        Solver(cfg, nx, ny)
