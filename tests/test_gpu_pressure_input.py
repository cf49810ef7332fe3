"""The +x pressure_input boundary (pressure_input_bc.f90) on the HIP path against the oracle (run on an MI355X).

FP64 results are bit-identical to the oracle except where a subsonic inlet row's isentropic relation calls
pow (the device's and the host C library's pow may differ in the last bit); the asserted bar is the
north-star tolerance, relative L-infinity <= 1e-12 per conserved variable, and rows without pow are checked
bit for bit."""
import sys

import numpy as np
import pytest

from fvleg_2d_amd import SchemeConfig, Solver
from fvleg_2d_amd._native import CatoError
from fvleg_2d_amd.problems import uniform_nodes

pytestmark = pytest.mark.gpu
TOL = 1e-12
TABLE = (np.array([0.0, 0.02, 0.05, 0.2]), np.array([1.8, 1.8, 0.6, 0.9]), np.array([1.3, 1.3, 0.5, 0.8]))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_linf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pad(a, nh):
    return np.asfortranarray(np.pad(np.asarray(a, dtype=np.float64), nh))


def drive_state(ni, nj, seed):
    """Near-rest gas with rows leaving and entering at +x (u of both signs, sub- and supersonic)."""
    rng = np.random.default_rng(seed)
    y = (np.arange(nj) + 0.5) / nj
    rho = np.full((ni, nj), 1.0) + 0.1 * rng.uniform(size=(ni, nj))
    u = np.tile(0.6 * np.sin(2 * np.pi * y), (ni, 1)) + 0.02 * rng.uniform(-1, 1, (ni, nj))
    u[-1, nj // 3] = 1.6  # supersonic outlet row
    u[-1, 2 * nj // 3] = -1.7  # supersonic inlet row
    v = 0.1 * rng.uniform(-1, 1, (ni, nj))
    p = np.full((ni, nj), 1.0) + 0.05 * rng.uniform(size=(ni, nj))
    return [np.asfortranarray(a) for a in (rho, u, v, p)]


def oracle_cfg(O, cfg, ni, nj, table=None, bc_time=0.0):
    return O.Config(ni, nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                    gamma=cfg.gamma, cfl=cfg.cfl, apply_low_mach_fix=cfg.apply_low_mach_fix,
                    bc_pressure=cfg.bc_pressure, bc_density=cfg.bc_density, pressure_table=table, bc_time=bc_time)


def inlet_rows(rho, u, p, gamma):
    """Rows whose last interior cell is a subsonic inlet (the pow branch)."""
    m = u / np.sqrt(gamma * p / rho)
    return (m <= 0.0) & (np.abs(m) <= 1.0)


@pytest.mark.parametrize("mode", ["constant", "table"])
@pytest.mark.parametrize("y", ["periodic", "symmetry", "zero_gradient"])
def test_apply_bc_and_residual_parity(oracle, mode, y):
    say("pressure_input apply_bc/residual", mode, y)
    ni, nj = 21, 17
    cfg = SchemeConfig(flux="SLAU", limiter="MLP5", integrator="ssp_rk3",
                       bcs=("pressure_input", "zero_gradient", y, y), bc_pressure=1.2, bc_density=0.9)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh)
    prims = drive_state(ni, nj, 4)
    t = 0.037  # between two knots of the table
    oc = oracle_cfg(oracle, cfg, ni, nj, TABLE if mode == "table" else None, bc_time=t)
    og = oracle.Geometry(oc, nx, ny)
    want = [pad(a, nh) for a in prims]
    oracle.apply_bcs(oc, *want)
    with Solver(cfg, nx, ny) as s:
        if mode == "table":
            s.set_pressure_input(TABLE)
        s.set_time(t)
        got = [pad(a, nh) for a in prims]
        s.apply_bc(*got)
        gd = s.residual(*[pad(a, nh) for a in prims])
    od = oracle.solve(oc, og, *[pad(a, nh) for a in prims])
    inl = inlet_rows(prims[0][-1], prims[1][-1], prims[3][-1], cfg.gamma)
    for k in range(4):
        assert rel_linf(got[k], want[k]) <= 1e-14, k
        # everything but the +x ghosts bit for bit, and those of the rows without pow
        assert np.array_equal(got[k][:nh + ni], want[k][:nh + ni]), k
        assert np.array_equal(got[k][nh + ni:, nh:nh + nj][:, ~inl], want[k][nh + ni:, nh:nh + nj][:, ~inl]), k
    for k in range(4):
        assert rel_linf(gd[k], od[k]) <= TOL, ("residual", k)
    assert inl.any() and (~inl).any()


def run_steps(oracle, cfg, nx, ny, prims, n, table, t0):
    ni, nj = prims[0].shape
    nh = cfg.n_ghost
    oc = oracle_cfg(oracle, cfg, ni, nj, table)
    f = oracle.Fluid(oc, oracle.Geometry(oc, nx, ny), *[pad(a, nh) for a in prims])
    f.set_time(t0)
    with Solver(cfg, nx, ny) as s:
        if table is not None:
            s.set_pressure_input(table)
        s.set_primitive(*prims)
        s.set_time(t0)
        for k in range(n):
            # one dt for both (the inlet rows' pow may move the CFL minimum by an ulp)
            dt = f.get_timestep()
            assert abs(s.compute_dt() - dt) <= 1e-12 * dt, k
            eg, eo = s.step(dt), f.integrate(dt)
            assert eg == eo == 0, (k, eg, eo)
        g = s.conserved()
    return g, [f.interior(nm) for nm in ("rho", "rho_u", "rho_v", "rho_E")], f.time


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3", "rk(1,3)", "rk(2,4)"])
@pytest.mark.parametrize("mode", ["constant", "table"])
def test_step_parity(oracle, integrator, mode):
    """Steps with host dt across the table's knots, from a start time t0 != 0 (the multistage schemes' first
    stage evaluates the table at t0 + dt: their U_prev copies a solver that never advanced)."""
    say("pressure_input steps", integrator, mode)
    ni, nj = 40, 24
    cfg = SchemeConfig(flux="AUSM+", limiter="minmod", integrator=integrator,
                       bcs=("pressure_input", "symmetry", "periodic", "periodic"), cfl=0.4,
                       bc_pressure=1.5, bc_density=1.1)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost, (0.0, 1.0), (0.0, 0.6))
    prims = drive_state(ni, nj, 7)
    g, o, t_end = run_steps(oracle, cfg, nx, ny, prims, 24, TABLE if mode == "table" else None, 0.011)
    assert t_end > TABLE[0][2]  # crossed two knots
    say("rel Linf", [rel_linf(g[k], o[k]) for k in range(4)])
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL, (k, rel_linf(g[k], o[k]))


@pytest.mark.parametrize("integrator", ["ssp_rk2", "ssp_rk3"])
def test_subcycled_steps_keep_the_solver_clock(oracle, integrator):
    """With subcycling every attempt advances the clock of the solver the first stage calls (U's own), so a
    retried step's first stage sees the table later than its other stages; host dt, table drive, vs the oracle."""
    say("pressure_input subcycle", integrator)
    ni, nj = 30, 12
    cfg = SchemeConfig(flux="AUSM+", limiter="minmod", integrator=integrator,
                       bcs=("pressure_input", "zero_gradient", "periodic", "periodic"), cfl=0.4)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh)
    prims = drive_state(ni, nj, 5)
    tol = 4e-3
    oc = oracle_cfg(oracle, cfg, ni, nj, TABLE)
    oc.allow_subcycle, oc.subcycle_tolerance = True, tol
    f = oracle.Fluid(oc, oracle.Geometry(oc, nx, ny), *[pad(a, nh) for a in prims])
    f.set_time(0.015)
    retried = 0
    with Solver(cfg, nx, ny) as s:
        s.set_pressure_input(TABLE)
        s.set_primitive(*prims)
        s.set_time(0.015)
        s.set_subcycle(True, tol)
        for k in range(10):
            dt = f.get_timestep()
            assert s.step(dt) == 0 and f.integrate(dt) == 0
            assert abs(s.last_dt() - f.last_dt) <= 1e-12 * f.last_dt, k
            retried += f.last_dt < dt
        g = s.conserved()
    o = [f.interior(nm) for nm in ("rho", "rho_u", "rho_v", "rho_E")]
    say("retried", retried, "rel Linf", [rel_linf(g[k], o[k]) for k in range(4)])
    assert retried > 0
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL, k


@pytest.mark.parametrize("integrator", ["ssp_rk3", "rk(2,4)"])
def test_device_loop_parity(oracle, integrator):
    """cato_run (device-side loop, 32-step graphs) against the oracle's main.f90 loop from a start time: the
    driver clock starts at the fluid time and each stage's boundary time comes from the device."""
    say("pressure_input device loop", integrator)
    ni, nj = 36, 20
    cfg = SchemeConfig(flux="SLAU", limiter="van_leer", integrator=integrator,
                       bcs=("pressure_input", "zero_gradient", "symmetry", "symmetry"), cfl=0.4,
                       bc_pressure=1.5, bc_density=1.1)
    nh = cfg.n_ghost
    nx, ny = uniform_nodes(ni, nj, nh)
    prims = drive_state(ni, nj, 9)
    t0, max_time = 0.004, 0.25
    oc = oracle_cfg(oracle, cfg, ni, nj, TABLE)
    f = oracle.Fluid(oc, oracle.Geometry(oc, nx, ny), *[pad(a, nh) for a in prims])
    f.set_time(t0)
    n_o, t_o, e_o = f.run(max_time, 2 ** 31 - 1, 0.05)
    with Solver(cfg, nx, ny) as s:
        s.set_pressure_input(TABLE)
        s.set_primitive(*prims)
        s.set_time(t0)
        n_g, t_g, e_g = s.run(max_time, 0, 0.05)
        g = s.conserved()
    o = [f.interior(nm) for nm in ("rho", "rho_u", "rho_v", "rho_E")]
    say("steps", n_g, n_o, "time", t_g, t_o, "rel Linf", [rel_linf(g[k], o[k]) for k in range(4)])
    assert e_g == e_o == 0 and n_g == n_o > 32 and abs(t_g - t_o) <= 1e-12 * t_o
    for k in range(4):
        assert rel_linf(g[k], o[k]) <= TOL, (k, rel_linf(g[k], o[k]))


def test_precision_form_close_to_fp64(oracle):
    say("pressure_input fp32_flux")
    ni, nj = 40, 24
    out = {}
    for prec in ("fp64", "fp32_flux"):
        cfg = SchemeConfig(flux="HLLC", limiter="MLP5", integrator="ssp_rk3",
                           bcs=("pressure_input", "symmetry", "periodic", "periodic"), cfl=0.4, precision=prec)
        nx, ny = uniform_nodes(ni, nj, cfg.n_ghost)
        with Solver(cfg, nx, ny) as s:
            s.set_pressure_input(TABLE)
            s.set_primitive(*drive_state(ni, nj, 2))
            s.advance(40)  # 32 of them in one graph
            out[prec] = s.conserved()
            assert np.all(np.isfinite(out[prec][0]))
    for k in range(4):
        a, b = out["fp32_flux"][k], out["fp64"][k]
        assert np.mean(np.abs(a - b)) / np.max(np.abs(b)) <= 1e-5, k


def test_errors():
    ni, nj = 8, 6
    nx, ny = uniform_nodes(ni, nj, 2)
    with pytest.raises(CatoError, match="pressure_bc location"):
        Solver(SchemeConfig(bcs=("zero_gradient", "pressure_input", "periodic", "periodic")), nx, ny)
    with pytest.raises(CatoError, match="outlet"):
        Solver(SchemeConfig(bcs=("pressure_input", "zero_gradient", "outflow", "symmetry"), bc_pressure=1.0), nx, ny)
    cfg = SchemeConfig(bcs=("pressure_input", "zero_gradient", "periodic", "periodic"))
    prims = drive_state(ni, nj, 1)
    with Solver(cfg, nx, ny) as s:
        s.set_primitive(*prims)
        with pytest.raises(CatoError, match="desired_pressure"):  # constants left at 0
            s.step(1e-3)
        with pytest.raises(CatoError, match="only one point"):
            s.set_pressure_input(([0.0], [1.0], [1.0]))
        with pytest.raises(CatoError, match="strictly increasing"):
            s.set_pressure_input(([0.0, 0.0], [1.0, 1.0], [1.0, 1.0]))
        s.set_pressure_input(([0.0, 1.0], [1.0, 1.0], [1.0, 1.0]))
        assert s.step(1e-3) == 0
        # a table whose pressure falls below tiny: NaN ghosts, reported as NaN by the step (error code 3)
        s.set_pressure_input(([0.0, 1.0], [1.0, -1.0], [1.0, 1.0]))
        s.set_time(0.9)
        assert s.step(1e-3) == 3
    from types import SimpleNamespace

    from fvleg_2d_amd.batch import Batch

    with Batch([SimpleNamespace(cfg=cfg, node_x=nx, node_y=ny)]) as bt:  # constants left at 0
        bt.set_primitive(0, *prims)
        with pytest.raises(CatoError, match="desired_pressure"):
            bt.run([0.01])
        with pytest.raises(CatoError, match="strictly increasing"):
            bt.set_pressure_input(0, ([0.0, 0.0], [1.0, 1.0], [1.0, 1.0]))


def drive_problem(ni, nj, seed, flux, limiter, bc_pressure, bc_density, y, precision):
    from types import SimpleNamespace

    cfg = SchemeConfig(flux=flux, limiter=limiter, integrator="ssp_rk3", bcs=("pressure_input", "zero_gradient", y, y),
                       cfl=0.4, bc_pressure=bc_pressure, bc_density=bc_density, precision=precision)
    nx, ny = uniform_nodes(ni, nj, cfg.n_ghost)
    rho, u, v, p = drive_state(ni, nj, seed)
    return SimpleNamespace(cfg=cfg, node_x=nx, node_y=ny, rho=rho, u=u, v=v, p=p)


def drive_mix(precision):
    """Three driven solves (a table, two constant drives; sub- and supersonic rows in and out) and one
    Riemann solve without a pressure input, all SSP-RK3."""
    import dataclasses

    from fvleg_2d_amd.problems import riemann2d

    plain = riemann2d(n=40, seed=2, flux="Roe", limiter="minmod", precision=precision)
    plain = dataclasses.replace(plain, max_time=0.1)
    return [drive_problem(36, 20, 9, "SLAU", "van_leer", 1.5, 1.1, "symmetry", precision),
            drive_problem(30, 24, 3, "HLLC", "MLP5", 0.8, 0.7, "periodic", precision),
            plain,
            drive_problem(28, 18, 5, "AUSMPW+", "minmod", 1.2, 0.9, "zero_gradient", precision)]


def alone(pb, table, max_time):
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        if table is not None:
            s.set_pressure_input(table)
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        n, t, e = s.run(max_time, 0, max_time)
        return s.conserved(), n, t, e


@pytest.mark.parametrize("precision", ["fp64", "fp32_flux"])
def test_batched_pressure_input_equals_solving_alone(precision):
    """The batched form's column passes (one workgroup per driven solve before every stage, in the captured
    step graph) give each solve the single-solve result bit for bit, a time table included."""
    from fvleg_2d_amd.batch import Batch

    say("batched pressure_input", precision)
    pbs = drive_mix(precision)
    tables = {0: TABLE}
    max_time = 0.1
    with Batch(pbs, precision=precision) as b:
        for k, pb in enumerate(pbs):
            b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
        b.set_pressure_input(0, TABLE)
        it, tt, er = b.run([max_time] * len(pbs), [max_time] * len(pbs))
        got = [b.conserved(k) for k in range(len(pbs))]
    for k, pb in enumerate(pbs):
        ref, n, t, e = alone(pb, tables.get(k), max_time)
        say("  solve", k, "steps", it[k], n)
        assert er[k] == e == 0 and it[k] == n > 20 and tt[k] == t
        for v in range(4):
            assert np.array_equal(got[k][v], ref[v]), (k, v)


@pytest.mark.parametrize("precision", ["fp64", "fp32_flux"])
def test_served_pressure_input_equals_solving_alone(precision):
    """Requests with a constant +x drive through the server, admitted while others run: each result is the
    single-solve result bit for bit."""
    from fvleg_2d_amd.serve import Server

    say("served pressure_input", precision)
    pbs = drive_mix(precision)[1:]  # the constant drives and the plain solve
    max_time = 0.08
    with Server(2e9, max_slots=4, precision=precision, steps_per_tick=8) as srv:
        ids = [srv.submit(pbs[0], max_time=max_time)]
        srv.tick()
        ids += [srv.submit(pb, max_time=max_time) for pb in pbs[1:]]
        srv.drain()
        got = [(srv.poll(i), srv.result(i)) for i in ids]
    for pb, ((done, _, n, e), res) in zip(pbs, got):
        ref, n1, _, e1 = alone(pb, None, max_time)
        assert done and e == e1 == 0 and n == n1 > 10
        for v in range(4):
            assert np.array_equal(res[v], ref[v]), v
