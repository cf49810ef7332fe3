"""Pins of the C oracle (CPU only).

The oracle is checked against (1) the reference's own golden vectors
(tests/golden/, from tests/unit/simple.h5 + mod_test_bc.f90 + mod_test_grid.f90),
(2) textbook values of the limiter functions, (3) invariants any correct
implementation satisfies exactly or to round-off (free stream, consistency
F(U,U) = F(U), mirror and transpose symmetry, conservation on a periodic
grid, linear exactness / no new extrema of the reconstruction), and (4) the
exact Riemann solution of the Sod problem (accuracy, not parity).
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT
from exact_riemann import sample

GOLD = os.path.join(ROOT, "tests", "golden")
FLUXES = ["AUSMPW+", "SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2", "Roe", "AUSM+", "HLLC", "M-AUSMPW+"]
LIMITERS = ["minmod", "superbee", "van_leer", "MLP3", "MLP5"]


def load_simple():
    d = json.load(open(os.path.join(GOLD, "simple_h5.json")))
    return {k: np.asfortranarray(np.array(d[k], dtype=np.float64)) for k in
            ("x", "y", "density", "x_velocity", "y_velocity", "pressure")}, d["n_ghost_layers"]


def fidx(i, nh=2):
    return i + nh - 1


@pytest.mark.parametrize("case", json.load(open(os.path.join(GOLD, "bc_expected.json")))["bc_tests"],
                         ids=lambda c: c["name"])
def test_bc_golden(oracle, case):
    d, nh = load_simple()
    b = case["deck_bcs"]
    cfg = oracle.Config(4, 4, bcs=(b["+x"], b["-x"], b["+y"], b["-y"]))
    f = {n: np.asfortranarray(d[k].copy()) for n, k in
         (("rho", "density"), ("u", "x_velocity"), ("v", "y_velocity"), ("p", "pressure"))}
    for a in f.values():  # zero_out_halo
        a[:nh, :] = 0.0
        a[-nh:, :] = 0.0
        a[:, :nh] = 0.0
        a[:, -nh:] = 0.0
    for bc_type, loc in case["apply"]:
        oracle.apply_bc_single(cfg, bc_type, loc, f["rho"], f["u"], f["v"], f["p"])
    for fld, kind, idx, (lo, hi), vals in case["checks"]:
        if kind == "col":
            got = f[fld][fidx(idx), fidx(lo):fidx(hi) + 1]
        else:
            got = f[fld][fidx(lo):fidx(hi) + 1, fidx(idx)]
        assert np.array_equal(got, np.array(vals, dtype=np.float64)), (fld, kind, idx, got, vals)


def test_geometry_golden(oracle):
    d, nh = load_simple()
    exp = json.load(open(os.path.join(GOLD, "bc_expected.json")))["grid"]
    for j in range(d["x"].shape[1]):
        assert np.array_equal(d["x"][:, j], exp["node_x_row"])
    for i in range(d["y"].shape[0]):
        assert np.array_equal(d["y"][i, :], exp["node_y_col"])
    cfg = oracle.Config(4, 4)
    g = oracle.Geometry(cfg, d["x"], d["y"])
    for e in range(4):
        assert np.all(g.edge_norm[0, e] == exp["edge_norm"][e][0])
        assert np.all(g.edge_norm[1, e] == exp["edge_norm"][e][1])
        assert np.all(g.edge_lengths[e] == exp["edge_lengths"][e])
    assert np.all(g.volume == exp["volume"])
    assert np.all(g.dx == 2.0) and np.all(g.dy == 1.0)


@pytest.mark.parametrize("lim,r,phi", [
    ("minmod", -1.0, 0.0), ("minmod", 0.5, 0.5), ("minmod", 2.0, 1.0),
    ("superbee", -1.0, 0.0), ("superbee", 0.25, 0.5), ("superbee", 0.75, 1.0), ("superbee", 1.5, 1.5),
    ("superbee", 3.0, 2.0),
    ("van_leer", -1.0, 0.0), ("van_leer", 0.5, 2.0 / 3.0), ("van_leer", 2.0, 4.0 / 3.0), ("van_leer", 1.0, 1.0),
])
def test_tvd_limiter_values(oracle, lim, r, phi):
    """Textbook limiter values (Sweby diagram): q = [.., 0, 1, 1+r, ..] gives the left face
    value q_i + phi(r)/2 with r = Delta_{i+1/2}/Delta_{i-1/2}."""
    cfg = oracle.Config(6, 2, limiter=lim)
    q = cfg.field()
    line = np.array([-2.0, -1.0, 0.0, 1.0, 1.0 + r, 1.0 + 2 * r, 1.0 + 3 * r, 1.0 + 4 * r, 1.0 + 5 * r, 1.0 + 6 * r])
    q[:, :] = line[:, None]
    ie, _ = oracle.interpolate_edges(cfg, q)
    # edge between halo'd index 3 (value 1) and 4 (value 1+r): cell i = 3 - nh + 1 = 2
    L = ie[0, 2, 1]
    assert abs((L - 1.0) - 0.5 * phi) < 1e-12


@pytest.mark.parametrize("lim", LIMITERS)
def test_reconstruction_constant_and_linear(oracle, lim):
    cfg = oracle.Config(12, 10, limiter=lim)
    q = cfg.field()
    q[:] = 3.25
    ie, je = oracle.interpolate_edges(cfg, q)
    assert np.all(ie == 3.25) and np.all(je == 3.25)
    # linear in i: face values are exact midpoints
    nx, ny = cfg.shape
    q[:] = (0.1 * np.arange(nx))[:, None] + 0.0 * np.arange(ny)[None, :]
    ie, _ = oracle.interpolate_edges(cfg, q)
    nh = cfg.nh
    for i in range(0, cfg.ni + 1):
        face = 0.1 * (i + nh - 1) + 0.05
        assert abs(ie[0, i, 3] - face) < 1e-12 and abs(ie[1, i, 3] - face) < 1e-12


@pytest.mark.parametrize("lim", LIMITERS)
def test_reconstruction_no_new_extrema(oracle, lim):
    rng = np.random.default_rng(3)
    cfg = oracle.Config(40, 6, limiter=lim)
    q = cfg.field()
    steps = rng.uniform(0.0, 1.0, cfg.shape[0])
    steps[rng.uniform(size=steps.size) < 0.3] *= 20.0
    q[:] = np.cumsum(steps)[:, None]
    ie, _ = oracle.interpolate_edges(cfg, q)
    nh = cfg.nh
    for i in range(0, cfg.ni + 1):
        a, b = q[i + nh - 1, 3], q[i + nh, 3]
        assert a - 1e-12 <= ie[0, i, 3] <= b + 1e-12
        assert a - 1e-12 <= ie[1, i, 3] <= b + 1e-12


def physical_flux(rho, u, v, p, g, nx, ny):
    E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    qn = u * nx + v * ny
    return np.array([rho * qn, rho * u * qn + p * nx, rho * v * qn + p * ny, qn * (E + p)])


@pytest.mark.parametrize("flux", FLUXES)
def test_flux_consistency(oracle, flux):
    """F(U, U) equals the physical Euler flux (uniform field, both directions)."""
    rng = np.random.default_rng(11)
    cfg = oracle.Config(4, 4, flux=flux)
    nodes = oracle.uniform_nodes(4, 4, 2)
    g = oracle.Geometry(cfg, *nodes)
    for _ in range(20):
        st = (rng.uniform(0.1, 2.0), rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.1, 1.5))
        edges = []
        for val in st:
            e = np.full((2, 5, 5), val, order="F")
            edges += [e, e]
        if flux == "M-AUSMPW+":
            edges = edges + edges  # superbee states of a uniform field are the same values
        p = cfg.field()
        p[:] = st[3]
        fi, fj = oracle.edge_fluxes(cfg, g, p, p, p, p, edges)
        for F, n in ((fi[:, 2, 1], (1.0, 0.0)), (fj[:, 1, 2], (0.0, 1.0))):
            ref = physical_flux(*st, 1.4, *n)
            assert np.allclose(F, ref, rtol=1e-13, atol=1e-13), (flux, st, F, ref)


def random_periodic(cfg, seed, amp=0.2):
    rng = np.random.default_rng(seed)
    nx, ny = cfg.shape
    x = np.linspace(0, 2 * np.pi, cfg.ni, endpoint=False)
    y = np.linspace(0, 2 * np.pi, cfg.nj, endpoint=False)
    X, Y = np.meshgrid(x, y, indexing="ij")
    fields = []
    for base, scale in ((1.0, 0.3), (0.0, 0.6), (0.0, 0.6), (1.0, 0.3)):
        a, b, c, d = rng.uniform(0.5, 1.5, 4)
        f = base + scale * (np.sin(a * X + b) * np.cos(c * Y + d)) + amp * rng.uniform(-0.05, 0.05, X.shape)
        full = cfg.field()
        full[cfg.nh:-cfg.nh, cfg.nh:-cfg.nh] = f
        fields.append(full)
    fields[0] = np.abs(fields[0]) + 0.2
    fields[3] = np.abs(fields[3]) + 0.2
    return fields


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("lim", ["minmod", "van_leer", "MLP5"])
def test_free_stream_preserved_exactly(oracle, flux, lim):
    cfg = oracle.Config(9, 7, flux=flux, limiter=lim)
    nodes = oracle.uniform_nodes(9, 7, cfg.nh, (0, 1.3), (0, 0.7))
    g = oracle.Geometry(cfg, *nodes)
    rho, u, v, p = (cfg.field() for _ in range(4))
    rho[:] = 1.7
    u[:] = 0.31
    v[:] = -0.42
    p[:] = 0.9
    d = oracle.solve(cfg, g, rho, u, v, p)
    for x in d:
        assert np.all(x == 0.0)


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("lim", LIMITERS)
def test_mirror_and_transpose_symmetry(oracle, flux, lim):
    """residual(mirror_x(U)) = mirror_x(residual(U)) with rho*u negated, and
    residual(transpose(U)) = transpose(residual(U)) with u <-> v: pins stencil
    orientation and the i/j specialisations against each other."""
    n = 12
    cfg = oracle.Config(n, n, flux=flux, limiter=lim, bcs=("periodic",) * 4)
    nodes = oracle.uniform_nodes(n, n, cfg.nh)
    g = oracle.Geometry(cfg, *nodes)
    rho, u, v, p = random_periodic(cfg, 5)
    d0 = oracle.solve(cfg, g, rho.copy(order="F"), u.copy(order="F"), v.copy(order="F"), p.copy(order="F"))
    nh = cfg.nh
    inner = (slice(nh, nh + n), slice(nh, nh + n))

    def mir(a):
        b = a.copy(order="F")
        b[inner] = a[inner][::-1, :]
        return b

    dm = oracle.solve(cfg, g, mir(rho), -mir(u), mir(v), mir(p))
    scale = max(np.abs(x[inner]).max() for x in d0)
    sign = [1, -1, 1, 1]
    for k in range(4):
        diff = np.abs(dm[k][inner] - sign[k] * d0[k][inner][::-1, :]).max()
        assert diff <= 1e-11 * scale, (k, diff, scale)

    def tr(a):
        b = a.copy(order="F")
        b[inner] = a[inner].T
        return b

    dt_ = oracle.solve(cfg, g, tr(rho), tr(v), tr(u), tr(p))
    perm = [0, 2, 1, 3]
    for k in range(4):
        diff = np.abs(dt_[k][inner] - d0[perm[k]][inner].T).max()
        assert diff <= 1e-11 * scale, (k, diff, scale)


@pytest.mark.parametrize("flux", FLUXES)
def test_conservation_on_periodic_grid(oracle, flux):
    cfg = oracle.Config(16, 12, flux=flux, limiter="minmod", bcs=("periodic",) * 4)
    nodes = oracle.uniform_nodes(16, 12, cfg.nh, (0, 1), (0, 0.75))
    g = oracle.Geometry(cfg, *nodes)
    rho, u, v, p = random_periodic(cfg, 9)
    d = oracle.solve(cfg, g, rho, u, v, p)
    nh = cfg.nh
    inner = (slice(nh, -nh), slice(nh, -nh))
    for x in d:
        tot = np.sum(x[inner] * g.volume[inner])
        mag = np.sum(np.abs(x[inner] * g.volume[inner]))
        assert abs(tot) <= 1e-13 * mag + 1e-300


@pytest.mark.parametrize("flux", ["AUSM+", "SLAU", "AUSMPW+", "Roe", "HLLC", "M-AUSMPW+"])
def test_sod_against_exact_riemann(oracle, flux):
    """Accuracy (not parity): density L1 error of Sod at t = 0.2 on 400 cells."""
    ni, nj = 400, 4
    cfg = oracle.Config(ni, nj, flux=flux, limiter="minmod", integrator="ssp_rk2",
                        bcs=("zero_gradient", "zero_gradient", "periodic", "periodic"), cfl=0.4)
    nx, ny = oracle.uniform_nodes(ni, nj, cfg.nh, (0, 1), (0, 0.01))
    g = oracle.Geometry(cfg, nx, ny)
    xc = 0.5 * (nx[:-1, :-1] + nx[1:, :-1])
    rho = np.asfortranarray(np.where(xc < 0.5, 1.0, 0.125))
    p = np.asfortranarray(np.where(xc < 0.5, 1.0, 0.1))
    z = cfg.field()
    f = oracle.Fluid(cfg, g, rho, z, z.copy(order="F"), p)
    n, t, err = f.run(0.2, contour_interval_dt=0.2)
    assert err == 0 and t == 0.2 and n > 100
    xi = xc[cfg.nh:-cfg.nh, cfg.nh]
    ex = sample(xi, 0.2, 0.5, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1)
    r = f.interior("rho")
    l1 = np.mean(np.abs(r[:, 0] - ex[0]))
    assert l1 < 0.012, l1
    assert np.abs(r - r[:, :1]).max() < 1e-8  # 1-D problem stays 1-D


def test_run_loop_terminates_on_clipped_time(oracle):
    """main.f90:178-194: the driver time lands on the contour cadence exactly and the loop
    stops there; the state advances by whole CFL steps."""
    ni, nj = 50, 4
    cfg = oracle.Config(ni, nj, flux="SLAU", limiter="minmod", integrator="ssp_rk2",
                        bcs=("zero_gradient", "zero_gradient", "periodic", "periodic"), cfl=0.4)
    nx, ny = oracle.uniform_nodes(ni, nj, cfg.nh, (0, 1), (0, 0.08))
    g = oracle.Geometry(cfg, nx, ny)
    xc = 0.5 * (nx[:-1, :-1] + nx[1:, :-1])
    rho = np.asfortranarray(np.where(xc < 0.5, 1.0, 0.125))
    p = np.asfortranarray(np.where(xc < 0.5, 1.0, 0.1))
    z = cfg.field()
    f = oracle.Fluid(cfg, g, rho, z, z.copy(order="F"), p)
    n, t, err = f.run(0.05, contour_interval_dt=0.01)
    assert err == 0
    assert t == 0.01 + 0.01 + 0.01 + 0.01 + 0.01 or abs(t - 0.05) < 1e-15
    f2 = oracle.Fluid(cfg, g, rho, z, z.copy(order="F"), p)
    n2 = 0
    tt = 0.0
    while tt < 0.05:  # without the clip: plain accumulation of CFL steps
        dt = f2.get_timestep()
        f2.integrate(dt)
        tt += dt
        n2 += 1
    assert n <= n2 + 5


def test_eos_closed_forms_and_mach_filter(oracle):
    """primitive -> conserved -> primitive (equation_of_state.f90:286-321, mod_fluid.f90:564-591) against the
    ideal-gas closed forms; the Mach filter zeroes |u/cs| < 1e-13 only."""
    import ctypes as C

    rng = np.random.default_rng(21)
    cfg = oracle.Config(6, 5, gamma=1.4)
    nodes = oracle.uniform_nodes(6, 5, cfg.nh)
    g = oracle.Geometry(cfg, *nodes)
    rho, u, v, p = (cfg.field() for _ in range(4))
    nh = cfg.nh
    inner = (slice(nh, -nh), slice(nh, -nh))
    rho[inner] = rng.uniform(0.1, 2.0, (6, 5))
    u[inner] = rng.uniform(-1.0, 1.0, (6, 5))
    v[inner] = rng.uniform(-1.0, 1.0, (6, 5))
    p[inner] = rng.uniform(0.1, 1.5, (6, 5))
    cs_true = np.sqrt(1.4 * p[inner] / rho[inner])
    u[nh, nh] = 1e-15 * cs_true[0, 0]  # |M| < 1e-13: filtered
    v[nh + 1, nh] = 5e-13 * cs_true[1, 0]  # |M| > 1e-13: kept
    f = oracle.Fluid(cfg, g, rho, u, v, p)
    E = p[inner] / 0.4 + 0.5 * rho[inner] * (u[inner] ** 2 + v[inner] ** 2)
    assert np.allclose(f.rho_E[inner], E, rtol=2e-15, atol=0)
    assert np.array_equal(f.rho_u[inner], rho[inner] * u[inner])
    out = [cfg.field() for _ in range(4)]
    c = cfg.c()
    oracle.lib().or_derived_quantities(C.byref(c), *[oracle._p(a) for a in (f.rho, f.rho_u, f.rho_v, f.rho_E)],
                                       *[oracle._p(a) for a in out])
    du, dv, dp, dcs = out
    assert du[nh, nh] == 0.0
    assert dv[nh + 1, nh] != 0.0
    mask = np.ones((6, 5), bool)
    mask[0, 0] = False
    assert np.allclose(du[inner][mask], u[inner][mask], rtol=1e-13, atol=1e-15)
    assert np.allclose(dv[inner], v[inner], rtol=1e-13, atol=1e-15)
    assert np.allclose(dp[inner], p[inner], rtol=1e-13)
    assert np.allclose(dcs[inner], np.sqrt(1.4 * dp[inner] / rho[inner]), rtol=1e-15)


def test_timestep_is_the_cfl_minimum(oracle):
    """get_timestep (mod_fluid.f90:554-558): min over cells of cfl / ((|u|+cs)/dx + (|v|+cs)/dy),
    brute force on a small non-uniform, non-square grid."""
    rng = np.random.default_rng(5)
    ni, nj = 7, 5
    cfg = oracle.Config(ni, nj, cfl=0.37, gamma=1.4)
    nodes = oracle.uniform_nodes(ni, nj, cfg.nh, (0.0, 1.3), (0.0, 0.4))
    g = oracle.Geometry(cfg, *nodes)
    rho, u, v, p = (cfg.field() for _ in range(4))
    nh = cfg.nh
    inner = (slice(nh, -nh), slice(nh, -nh))
    rho[inner] = rng.uniform(0.2, 2.0, (ni, nj))
    u[inner] = rng.uniform(-2.0, 2.0, (ni, nj))
    v[inner] = rng.uniform(-2.0, 2.0, (ni, nj))
    p[inner] = rng.uniform(0.1, 1.5, (ni, nj))
    f = oracle.Fluid(cfg, g, rho, u, v, p)
    dx, dy = 1.3 / ni, 0.4 / nj
    cs = np.sqrt(1.4 * p[inner] / rho[inner])
    per_cell = 0.37 / ((np.abs(u[inner]) + cs) / dx + (np.abs(v[inner]) + cs) / dy)
    assert abs(f.get_timestep() - per_cell.min()) <= 1e-14 * per_cell.min()


def test_m_ausmpw_strong_shock_tube_against_exact_riemann(oracle):
    """M-AUSMPW+ blends in the superbee states where a jump exceeds 1 (phi_L_half / phi_R_half,
    m_ausmpw_plus_solver.fypp:790-827): a shock tube with a pressure ratio of 100 exercises that branch;
    density L1 error against the exact solution at t = 0.012 on 400 cells, and the 1-D invariance."""
    ni, nj = 400, 4
    cfg = oracle.Config(ni, nj, flux="M-AUSMPW+", limiter="minmod", integrator="ssp_rk2",
                        bcs=("zero_gradient", "zero_gradient", "periodic", "periodic"), cfl=0.3)
    nx, ny = oracle.uniform_nodes(ni, nj, cfg.nh, (0, 1), (0, 0.01))
    g = oracle.Geometry(cfg, nx, ny)
    xc = 0.5 * (nx[:-1, :-1] + nx[1:, :-1])
    rho = np.asfortranarray(np.where(xc < 0.5, 1.0, 0.125))
    p = np.asfortranarray(np.where(xc < 0.5, 10.0, 0.1))
    z = cfg.field()
    f = oracle.Fluid(cfg, g, rho, z, z.copy(order="F"), p)
    n, t, err = f.run(0.012, contour_interval_dt=0.012)
    assert err == 0 and n > 50
    xi = xc[cfg.nh:-cfg.nh, cfg.nh]
    ex = sample(xi, t, 0.5, 1.0, 0.0, 10.0, 0.125, 0.0, 0.1)
    r = f.interior("rho")
    l1 = np.mean(np.abs(r[:, 0] - ex[0]))
    assert l1 < 0.03, l1
    assert np.abs(r - r[:, :1]).max() < 1e-8


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(2,4)"])
def test_energy_source_adds_s_dt_to_rho_E(oracle, integrator):
    """mod_fluid.f90:500-509: d(rho E)/dt gains the source field in every stage, so on a uniform periodic
    state one step of any consistent scheme adds S*dt to rho E and leaves rho, rho u, rho v alone."""
    n = 8
    cfg = oracle.Config(n, n, flux="SLAU", limiter="minmod", integrator=integrator, bcs=("periodic",) * 4)
    g = oracle.Geometry(cfg, *oracle.uniform_nodes(n, n, cfg.nh))
    rho, u, v, p = (cfg.field() for _ in range(4))
    rho[:] = 1.3
    u[:] = 0.2
    v[:] = -0.1
    p[:] = 0.9
    f = oracle.Fluid(cfg, g, rho, u, v, p)
    before = {k: f.interior(k) for k in ("rho", "rho_u", "rho_v", "rho_E")}
    S = 0.37
    f.set_energy_source(np.full(cfg.shape, S, order="F"))
    dt = 0.01
    assert f.integrate(dt) == 0
    assert np.allclose(f.interior("rho_E") - before["rho_E"], S * dt, rtol=0, atol=1e-13)
    for k in ("rho", "rho_u", "rho_v"):
        assert np.allclose(f.interior(k), before[k], rtol=1e-15, atol=0)


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(2,4)"])
def test_convergence_norms_closed_form(oracle, integrator):
    """calculate_convergence (mod_fluid.f90:1162-1205) on a uniform periodic state heated by a constant source S:
    every stage of SSP-RK2/RK3 after the first lands on U + S dt again, so the norm of (last - first stage) is
    round-off; rk_multistage compares against U_0 (mod_fluid.f90:970), giving S dt / (rho E) exactly in rho E."""
    n = 12
    cfg = oracle.Config(n, n, flux="AUSMPW+", limiter="van_leer", integrator=integrator, bcs=("periodic",) * 4)
    g = oracle.Geometry(cfg, *oracle.uniform_nodes(n, n, cfg.nh))
    rho, u, v, p = (cfg.field() for _ in range(4))
    rho[:], u[:], v[:], p[:] = 0.8, 0.0, 0.0, 1.1
    f = oracle.Fluid(cfg, g, rho, u, v, p)
    rho_E0 = f.interior("rho_E")[0, 0]
    S, dt = 2.5, 0.004
    f.set_energy_source(np.full(cfg.shape, S, order="F"))
    assert f.integrate(dt) == 0
    conv = f.convergence()
    assert conv[:3] == [0.0, 0.0, 0.0]
    if integrator == "rk(2,4)":
        assert abs(conv[3] - S * dt / rho_E0) <= 1e-12 * conv[3]
    else:
        assert conv[3] < 1e-14


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(1,3)"])
def test_convergence_norms_against_numpy(oracle, integrator):
    """The same norms from numpy.linalg.norm on the states themselves: U_1 = U + L(U) dt built from the oracle's
    solve (flushes at 2.2e-16 change nothing at this tolerance), the momentum norms divided by max(norm, 1)."""
    ni, nj = 24, 17
    cfg = oracle.Config(ni, nj, flux="SLAU", limiter="MLP3", integrator=integrator, bcs=("zero_gradient",) * 4)
    g = oracle.Geometry(cfg, *oracle.uniform_nodes(ni, nj, cfg.nh, (0.0, 1.0), (0.0, 0.7)))
    rng = np.random.default_rng(3)
    rho, u, v, p = (cfg.field() for _ in range(4))
    rho[:] = rng.uniform(0.5, 2.0, rho.shape)
    u[:] = rng.uniform(-0.4, 0.4, u.shape)
    v[:] = rng.uniform(-0.01, 0.01, v.shape)  # norm2(rho v) < 1: the max(., 1) branch; rho u takes the other
    p[:] = rng.uniform(0.5, 1.5, p.shape)
    f = oracle.Fluid(cfg, g, rho, u, v, p)
    U0 = [f.interior(k) for k in ("rho", "rho_u", "rho_v", "rho_E")]
    L0 = oracle.solve(cfg, g, *[getattr(f, k).copy(order="F") for k in ("rho", "u", "v", "p")])
    nh = cfg.nh
    dt = 0.5 * f.get_timestep()
    assert f.integrate(dt) == 0
    U = [f.interior(k) for k in ("rho", "rho_u", "rho_v", "rho_E")]
    first = U0 if integrator.startswith("rk(") else [a + b[nh:-nh, nh:-nh] * dt for a, b in zip(U0, L0)]
    want = []
    for k in range(4):
        den = np.linalg.norm(first[k])
        want.append(np.linalg.norm(U[k] - first[k]) / (max(den, 1.0) if k in (1, 2) else den))
    got = f.convergence()
    assert np.linalg.norm(first[2]) < 1.0 < np.linalg.norm(first[1])
    assert min(want) > 1e-6
    assert np.allclose(got, want, rtol=1e-12, atol=0), (got, want)


def outflow_setup(oracle, bcs, seed=5, ni=9, nj=7):
    cfg = oracle.Config(ni, nj, flux="SLAU", limiter="minmod", bcs=bcs, bc_pressure=0.8)
    rng = np.random.default_rng(seed)
    nh = cfg.nh
    rho, u, v, p = (cfg.field() for _ in range(4))
    sl = (slice(nh, nh + ni), slice(nh, nh + nj))
    rho[sl] = rng.uniform(0.5, 2.0, (ni, nj))
    u[sl] = rng.uniform(-2.0, 2.0, (ni, nj))  # some rows supersonic at +x
    v[sl] = rng.uniform(-1.0, 1.0, (ni, nj))
    p[sl] = rng.uniform(0.3, 1.5, (ni, nj))
    p[nh + ni - 1, nh + 1] = 0.8  # one row exactly at the ambient pressure
    return cfg, [rho, u, v, p]


def test_outflow_bc_characteristics(oracle):
    """outlet_bc.f90:96-123 at +x, checked against what the characteristic theory fixes (Blazek 2015, 8.4):
    a supersonic row is copied; a subsonic row gets p = p_b and keeps the outgoing acoustic invariant
    p + rho c u_n, the entropy invariant p - c^2 rho and the tangential velocity of the last interior cell
    (c, n from that cell), the same in every ghost layer; a row at p_b is copied exactly."""
    cfg, f = outflow_setup(oracle, ("outflow", "zero_gradient", "zero_gradient", "zero_gradient"))
    rho, u, v, p = f
    oracle.apply_bcs(cfg, rho, u, v, p)
    nh, ni, nj, pb, gam = cfg.nh, cfg.ni, cfg.nj, cfg.bc_pressure, cfg.gamma
    i_in = nh + ni - 1
    n_sub = n_sup = 0
    for j in range(nj + 2 * nh):
        d = [a[i_in, j] for a in (rho, u, v, p)]
        c = np.sqrt(gam * d[3] / d[0])
        vel = np.hypot(d[1], d[2])
        for g in range(1, nh + 1):
            gh = [a[i_in + g, j] for a in (rho, u, v, p)]
            if vel / c > 1.0 or d[3] == pb:
                assert gh == d
                n_sup += vel / c > 1.0
                continue
            n_sub += 1
            nx, ny = d[1] / vel, d[2] / vel
            assert gh[3] == pb
            jp = lambda q: q[3] + d[0] * c * (q[1] * nx + q[2] * ny)
            assert abs(jp(gh) - jp(d)) <= 1e-13 * abs(jp(d))
            assert abs((gh[3] - c * c * gh[0]) - (d[3] - c * c * d[0])) <= 1e-13 * c * c * d[0]
            assert abs((gh[1] * ny - gh[2] * nx) - (d[1] * ny - d[2] * nx)) <= 1e-13 * vel
    assert n_sub > 0 and n_sup > 0


def test_outflow_bc_priority_and_other_sides(oracle):
    """Priority 0 (outlet_bc.f90:54): the y fills of higher priority run first, so the +x corners take the
    extrapolation of the y-filled last column (a mirrored row under symmetry); -x, +y, -y outflow copy the
    last interior line over the whole extent, after the others."""
    cfg, f = outflow_setup(oracle, ("outflow", "outflow", "symmetry", "zero_gradient"))
    ref = [a.copy(order="F") for a in f]
    rho, u, v, p = f
    oracle.apply_bcs(cfg, rho, u, v, p)
    nh, ni, nj = cfg.nh, cfg.ni, cfg.nj
    # -x: plain copy of column 1 over the whole j extent (incl. the y ghosts filled first)
    for a in f:
        assert np.array_equal(a[:nh, :], np.repeat(a[nh:nh + 1, :], nh, axis=0))
    # +x corners at +y (symmetry): the extrapolation of the mirrored cell, v negated before the fill
    one = oracle.Config(ni, nj, flux="SLAU", limiter="minmod", bcs=("outflow", "outflow", "zero_gradient",
                        "zero_gradient"), bc_pressure=0.8)
    m = [a.copy(order="F") for a in ref]
    for g in range(1, nh + 1):  # symmetry at +y: row nj+g mirrors row nj+1-g, v negated
        for k, a in enumerate(m):
            a[:, nh + nj - 1 + g] = (-1 if k == 2 else 1) * a[:, nh + nj - g]
    oracle.apply_bc_single(one, "outflow", "+x", *m)
    for k in range(4):
        assert np.array_equal(f[k][nh + ni:, nh + nj:], m[k][nh + ni:, nh + nj:]), k


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(2,4)"])
def test_subcycle_retries_with_a_tenth_of_dt(oracle, integrator):
    """subcycle_check (mod_fluid.f90:856-887): with a tolerance between the convergence norms of the step at dt and
    at dt/10, integrate(dt) retries once, reports dt/10 and lands bit for bit on a plain integrate(dt/10) of the same
    state (the retry starts again from U); a tolerance above the first norms changes nothing; rk_multistage has no
    retry loop."""
    ni, nj = 24, 16
    rng = np.random.default_rng(17)
    base = dict(flux="SLAU", limiter="minmod", integrator=integrator,
                bcs=("zero_gradient", "zero_gradient", "periodic", "periodic"), cfl=0.4)
    prims = [rng.uniform(0.5, 1.5, (ni, nj)), rng.uniform(-0.5, 0.5, (ni, nj)), rng.uniform(-0.5, 0.5, (ni, nj)),
             rng.uniform(0.5, 1.5, (ni, nj))]

    def fluid(**kw):
        cfg = oracle.Config(ni, nj, **base, **kw)
        nx, ny = oracle.uniform_nodes(ni, nj, cfg.nh)
        return oracle.Fluid(cfg, oracle.Geometry(cfg, nx, ny), *[np.pad(a, cfg.nh) for a in prims])

    def state(f):
        return [f.interior(n) for n in ("rho", "rho_u", "rho_v", "rho_E")]

    fa = fluid()
    dt = fa.get_timestep()
    assert fa.integrate(dt) == 0 and fa.last_dt == dt
    n1 = max(fa.convergence())
    fb = fluid()
    assert fb.integrate(dt / 10.0) == 0
    n2 = max(fb.convergence())
    assert n2 < 0.5 * n1  # the norms shrink with dt
    tol = 0.5 * (n1 + n2)
    fc = fluid(allow_subcycle=True, subcycle_tolerance=tol)
    assert fc.integrate(dt) == 0
    retried = integrator != "rk(2,4)"
    assert fc.last_dt == (dt / 10.0 if retried else dt)
    assert fc.time == fc.last_dt
    want = fb if retried else fa
    for a, b in zip(state(fc), state(want)):
        assert np.array_equal(a, b)
    assert fc.convergence() == want.convergence()
    fd = fluid(allow_subcycle=True, subcycle_tolerance=2.0 * n1)
    assert fd.integrate(dt) == 0 and fd.last_dt == dt
    for a, b in zip(state(fd), state(fa)):
        assert np.array_equal(a, b)
