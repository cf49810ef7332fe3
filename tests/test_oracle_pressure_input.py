"""Pins of the oracle's +x pressure_input boundary (pressure_input_bc.f90, inlet_outlet.f90) and of its
pressure table (finterp's linear_interp_1d), against what the mathematics fixes: knots, lines and
numpy.interp for the table; the characteristic and isentropic relations of Blazek (2015, 8.4-8.5) for the
inlet and outlet states; the branch rules of pressure_input_bc.f90:313-341 on hand-made columns.

The boundary time of each stage (the flux solvers' own clocks, mod_fluid.f90) has no independent check:
its parity is unpinned beyond the reading cited in the oracle."""
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "pressure_input")
GAMMA = 1.4


def test_table_interpolation_knots_lines_and_extrapolation(oracle):
    t = np.array([0.0, 0.5, 1.25, 2.0, 4.0])
    f = np.array([3.0, -1.0, 2.5, 2.5, 10.0])
    for k in range(len(t)):
        assert oracle.interp_linear(t, f, t[k]) == f[k]  # knots exactly, the last one included
    xs = np.linspace(0.0, 4.0, 97)
    got = np.array([oracle.interp_linear(t, f, x) for x in xs])
    assert np.allclose(got, np.interp(xs, t, f), rtol=0, atol=1e-14)
    # beyond the ends: the end intervals' lines
    for x in (-1.0, -0.25, 4.5, 9.0):
        k = 0 if x < t[0] else len(t) - 2
        line = f[k] + (f[k + 1] - f[k]) * (x - t[k]) / (t[k + 1] - t[k])
        assert abs(oracle.interp_linear(t, f, x) - line) <= 1e-13 * max(1.0, abs(line))
    # two knots: a single line everywhere
    assert oracle.interp_linear([1.0, 3.0], [2.0, 6.0], 2.0) == 4.0


def test_reader_on_reference_decks():
    from fvleg_2d_amd.pressure_input import read_pressure_input

    t, p, r = read_pressure_input(os.path.join(GOLD, "planar_target_2d.dat"))
    assert list(t) == [0.0, 0.060e-09, 0.061e-09, 1.0]
    assert list(p) == [3.5e13, 3.5e13, 1.0e9, 1.0e9] and list(r) == [0.001] * 4
    t, p, r = read_pressure_input(os.path.join(GOLD, "double_picket.dat"), scale_factor=2.0)
    assert len(t) == 8 and t[-1] == 1.0 and np.all(np.diff(t) > 0)
    assert p[4] == 2.0 * 1.0e11 and r[2] == 2.0 * 0.01  # the scale factor multiplies density too (:163-164)
    t, p, r = read_pressure_input(os.path.join(GOLD, "supported_shock.dat"))
    assert list(p) == [7.5e12, 7.5e12, 5.0e13, 5.0e13, 1.0e6]
    with pytest.raises(ValueError):  # time and pressure only: the reference's table fails initialize_1d
        read_pressure_input(os.path.join(GOLD, "two_column.dat"))


def pi_column(oracle, states, pd=1.0, rd=0.5, table=None, bc_time=0.0, y="zero_gradient"):
    """A grid whose column ni holds the given (rho, u, p) rows; returns cfg and the filled halo'd fields."""
    nj = len(states)
    cfg = oracle.Config(4, nj, flux="SLAU", limiter="minmod", bcs=("pressure_input", "zero_gradient", y, y),
                        gamma=GAMMA, bc_pressure=pd, bc_density=rd, pressure_table=table, bc_time=bc_time)
    nh = cfg.nh
    rho, u, v, p = (cfg.field() for _ in range(4))
    for a, val in ((rho, 1.0), (u, 0.1), (v, 0.3), (p, 1.0)):
        a[nh:nh + 4, nh:nh + nj] = val
    for j, (r_, u_, p_) in enumerate(states):
        rho[nh + 3, nh + j], u[nh + 3, nh + j], p[nh + 3, nh + j] = r_, u_, p_
        v[nh + 3, nh + j] = 0.7  # ignored by the boundary (domain v = 0, :312)
    oracle.apply_bcs(cfg, rho, u, v, p)
    return cfg, (rho, u, v, p)


def ghost(cfg, f, j, layer=1):
    nh, ni = cfg.nh, cfg.ni
    return [a[nh + ni - 1 + layer, nh + j] for a in f]


def test_outlet_rows(oracle):
    """u > 0: subsonic rows, and supersonic rows while the column's largest pressure is below the desired one,
    take subsonic_outlet: p = p_d exactly, the outgoing acoustic invariant p + rho c u and the entropy
    invariant p - c^2 rho of the domain cell kept (c of the domain cell); v = 0 in every ghost layer."""
    states = [(1.0, 0.3, 0.8), (0.5, 0.9, 0.4), (2.0, 1.5, 0.6)]  # M = 0.28, 0.87 (sub), 1.42 (super)
    cfg, f = pi_column(oracle, states, pd=0.7)  # the column's largest pressure (0.8) is above p_d
    for j, (r_, u_, p_) in enumerate(states):
        c = np.sqrt(GAMMA * p_ / r_)
        for layer in (1, 2):
            g = ghost(cfg, f, j, layer)
            assert g[2] == 0.0 and np.signbit(g[2]) == False  # noqa: E712
            if u_ / c > 1.0:
                assert g == [r_, u_, 0.0, p_]
                continue
            assert g[3] == 0.7
            assert abs((g[3] + r_ * c * g[1]) - (p_ + r_ * c * u_)) <= 1e-14
            assert abs((g[3] - c * c * g[0]) - (p_ - c * c * r_)) <= 1e-14
    # the same supersonic row with a desired pressure above the whole column: subsonic_outlet
    cfg, f = pi_column(oracle, states, pd=1.7)
    r_, u_, p_ = states[2]
    c = np.sqrt(GAMMA * p_ / r_)
    g = ghost(cfg, f, 2)
    assert g[3] == 1.7 and abs((g[3] + r_ * c * g[1]) - (p_ + r_ * c * u_)) <= 1e-14


def test_supersonic_and_subsonic_inlet_rows(oracle):
    """u <= 0: |M| > 1 takes (rho_d, u, 0, p_d) of the desired state; otherwise subsonic_inlet
    (inlet_outlet.f90:26-140), checked against Blazek's relations: the boundary keeps the domain's normal
    velocity (R- and R+ built from the same c_b), p and rho follow the total state isentropically,
    p_b/p_0 = (1 + (g-1)/2 M^2)^(-g/(g-1)) and rho_b/rho_0 = (p_b/p_0)^(1/g), and c_b solves the energy
    equation c_0^2 = c_b^2 + (g-1)/2 q_b^2 with q_b = (R- + 2 c_b/(g-1)) / cos(theta), cos(theta) = -1."""
    pd, rd = 1.6, 0.9
    states = [(1.0, -2.0, 1.0), (1.0, -0.4, 1.0), (0.7, -0.05, 0.5), (1.2, 0.0, 0.9)]
    cfg, f = pi_column(oracle, states, pd=pd, rd=rd)
    g = ghost(cfg, f, 0)
    assert g == [rd, -2.0, 0.0, pd]
    for j in (1, 2, 3):
        r_, u_, p_ = states[j]
        g = ghost(cfg, f, j)
        c = np.sqrt(GAMMA * p_ / r_)
        assert abs(g[1] - u_) <= 1e-14 * c  # v_tot_b = 2 c_b/(g-1) - R+ = -v.n = u
        ratio = g[3] / pd
        assert abs(g[0] / rd - ratio ** (1.0 / GAMMA)) <= 1e-14
        if u_ == 0.0:  # cos(theta) = 0: c_b = -R- (g-1)/2 = c, M_b = 0: the total state itself
            assert g[0] == rd and g[3] == pd
            continue
        m2 = (ratio ** (-(GAMMA - 1.0) / GAMMA) - 1.0) * 2.0 / (GAMMA - 1.0)
        cb = abs(g[1]) / np.sqrt(m2)
        c0sq = c * c + 0.5 * (GAMMA - 1.0) * u_ * u_
        rminus = -u_ - 2.0 * c / (GAMMA - 1.0)
        qb = (rminus + 2.0 * cb / (GAMMA - 1.0)) / -1.0
        assert abs(c0sq - (cb * cb + 0.5 * (GAMMA - 1.0) * qb * qb)) <= 1e-9 * c0sq


def test_table_drives_the_desired_state_and_tiny_is_an_error(oracle):
    t = np.array([0.0, 1.0, 2.0])
    p = np.array([1.0, 3.0, 2.0])
    r = np.array([0.5, 0.25, 0.75])
    states = [(1.0, 0.3, 0.8)]  # a subsonic outlet row: ghost p = p_d(t)
    for bt, want in ((0.0, 1.0), (0.5, 2.0), (1.5, 2.5), (3.0, 1.0)):
        cfg, f = pi_column(oracle, states, table=(t, p, r), bc_time=bt)
        assert ghost(cfg, f, 0)[3] == oracle.interp_linear(t, p, bt)
        assert abs(ghost(cfg, f, 0)[3] - want) <= 1e-15
    # the table, not the constants, when both are set; a desired pressure below tiny fills NaN (the
    # reference stops with an error; the step then reports NaN)
    cfg, f = pi_column(oracle, states, pd=5.0, rd=5.0, table=(t, [1.0, 0.0, 0.0], r), bc_time=1.5)
    assert np.isnan(ghost(cfg, f, 0)[3])
    cfg, f = pi_column(oracle, states, pd=0.0, rd=1.0)
    assert np.isnan(ghost(cfg, f, 0)[0])


def test_corners_and_priority(oracle):
    """Priority 0 (pressure_input_bc.f90:78): the y sides fill first, so a corner ghost is the boundary state
    of the y-filled cell of column ni, i.e. of its mapped interior row; v = +0 there too."""
    states = [(1.0, 0.3, 0.8), (0.5, -0.9, 0.4), (2.0, 1.5, 0.6)]
    for y in ("zero_gradient", "symmetry", "periodic"):
        cfg, f = pi_column(oracle, states, pd=1.2, rd=0.6, y=y)
        nh, ni, nj = cfg.nh, cfg.ni, cfg.nj
        for g in range(1, nh + 1):
            top = {"zero_gradient": nj - 1, "symmetry": nj - g, "periodic": g - 1}[y]
            bot = {"zero_gradient": 0, "symmetry": g - 1, "periodic": nj - g}[y]
            for layer in (1, 2):
                assert ghost(cfg, f, nj - 1 + g, layer) == ghost(cfg, f, top, layer)
                assert ghost(cfg, f, -g, layer) == ghost(cfg, f, bot, layer)


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(2,4)"])
def test_fluid_time_and_constant_table(oracle, integrator):
    """The fluid time advances by dt per step from set_time; a table that is constant in time gives the
    constant boundary's results bit for bit (whatever time each stage evaluates it at: with powers of two,
    p1*f + q1*f = f exactly)."""
    from fvleg_2d_amd.problems import uniform_nodes

    ni, nj = 12, 6
    base = dict(flux="AUSM+", limiter="minmod", integrator=integrator,
                bcs=("pressure_input", "zero_gradient", "periodic", "periodic"), cfl=0.3)
    rng = np.random.default_rng(3)
    prims = [rng.uniform(0.6, 1.2, (ni, nj)), rng.uniform(-0.3, 0.3, (ni, nj)), rng.uniform(-0.2, 0.2, (ni, nj)),
             rng.uniform(0.6, 1.2, (ni, nj))]
    out = []
    for table in (None, ([0.0, 0.3, 2.0], [1.0, 1.0, 1.0], [0.5, 0.5, 0.5])):
        cfg = oracle.Config(ni, nj, bc_pressure=1.0, bc_density=0.5, pressure_table=table, **base)
        nx, ny = uniform_nodes(ni, nj, cfg.nh)
        g = oracle.Geometry(cfg, nx, ny)
        fl = oracle.Fluid(cfg, g, *[np.asfortranarray(np.pad(a, cfg.nh)) for a in prims])
        fl.set_time(0.25)
        t = 0.25
        for _ in range(4):
            dt = fl.get_timestep()
            assert fl.integrate(dt) == 0
            t = t + dt
            assert fl.time == t
        out.append([fl.interior(n) for n in ("rho", "rho_u", "rho_v", "rho_E")])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
