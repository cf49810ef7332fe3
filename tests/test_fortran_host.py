"""The Fortran host (fvleg_2d_amd/fortran): the reference's input.ini + pycato
HDF5 deck, the plugin objects over the C ABI, and the main.f90 time loop.

CPU: the host is built (by __graft_entry__.build()) and reads a deck without a
device (--dry-run).  GPU: the Sod shock tube of the reference's sod_1d deck
family runs through the host and matches the oracle bit for bit, with the
host-side loop and with the device-side loop.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HOST = os.path.join(ROOT, "fvleg_2d_amd", "fortran", "cato_hip_host")


def host_or_fail():
    if not os.path.exists(HOST):
        pytest.fail(f"{HOST} is not built; run __graft_entry__.build()")
    return HOST


def sod_problem(**kw):
    from fvleg_2d_amd.problems import sod

    return sod(**kw)


def test_host_reads_deck_without_device(tmp_path):
    from fvleg_2d_amd.deck import write_deck

    pb = sod_problem(ni=400, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk2", cfl=0.4, max_time=0.2)
    ini = write_deck(pb, str(tmp_path), "sod")
    r = subprocess.run([host_or_fail(), ini, "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert "ni = 400, nj = 4, n_ghost_layers = 2" in out
    # flux AUSM+ (6), minmod (0), ssp_rk2 (0), bcs zero_gradient x (1 1), periodic y (0 0), 2 ghost layers
    cfg_line = [l for l in out.splitlines() if l.startswith("config:")][0]
    assert cfg_line.split("=")[1].split() == ["6", "0", "0", "1", "1", "0", "0", "2"]
    vals = [l for l in out.splitlines() if l.startswith("gamma cfl max_time")][0].split("=")[1].split()
    assert [float(v) for v in vals] == [1.4, 0.4, 0.2]
    nodes = [l for l in out.splitlines() if l.startswith("x(1-nh")][0].split("=")[1].split()
    assert float(nodes[0]) == pb.node_x[0, 0] and float(nodes[1]) == pb.node_y[-1, -1]
    cells = [l for l in out.splitlines() if l.startswith("density(1,1)")][0].split("=")[1].split()
    assert float(cells[0]) == pb.rho[0, 0] and float(cells[1]) == pb.p[-1, -1]


def test_host_deck_defaults_follow_the_reference(tmp_path):
    """Keys a deck may omit take mod_input.f90's defaults: flux_solver 'AUSMPW+' (the component default,
    mod_input.f90:132, read without one at :264; the reference's noh, rayleigh_taylor and richtmyer_meshkov
    decks omit it) and contour_interval_dt 0.1 (mod_input.f90:436), which drives the driver-time clip."""
    from fvleg_2d_amd.deck import write_deck

    pb = sod_problem(ni=40, nj=4)
    ini = write_deck(pb, str(tmp_path), "sod")
    lines = [l for l in open(ini).read().splitlines()
             if not l.startswith("flux_solver") and not l.startswith("contour_interval_dt")]
    open(ini, "w").write("\n".join(lines) + "\n")
    r = subprocess.run([host_or_fail(), ini, "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    cfg_line = [l for l in r.stdout.splitlines() if l.startswith("config:")][0]
    assert cfg_line.split("=")[1].split()[0] == "0"  # AUSMPW+
    assert float([l for l in r.stdout.splitlines() if l.startswith("contour_interval_dt")][0].split("=")[1]) == 0.1


def test_host_rejects_unsupported_scheme(tmp_path):
    from fvleg_2d_amd.deck import write_deck

    pb = sod_problem(ni=40, nj=4)
    ini = write_deck(pb, str(tmp_path), "sod")
    text = open(ini).read().replace("flux_solver = 'AUSM+'", "flux_solver = 'FVLEG'")
    open(ini, "w").write(text)
    r = subprocess.run([host_or_fail(), ini, "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode != 0
    assert "unsupported flux_solver 'FVLEG'" in r.stdout


def test_host_reads_outflow_deck(tmp_path):
    """'outflow' at +x with its outflow_ambient_pressure (mod_input.f90:384-390, bc_factory.f90:53)."""
    from fvleg_2d_amd.deck import write_deck

    pb = sod_problem(ni=40, nj=4)
    pb.cfg.bcs = ("outflow", "zero_gradient", "periodic", "periodic")
    pb.cfg.bc_pressure = 0.25
    ini = write_deck(pb, str(tmp_path), "sod")
    r = subprocess.run([host_or_fail(), ini, "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    cfg_line = [l for l in r.stdout.splitlines() if l.startswith("config:")][0]
    assert cfg_line.split("=")[1].split()[3:7] == ["3", "1", "0", "0"]
    assert float([l for l in r.stdout.splitlines() if l.startswith("bc_pressure")][0].split("=")[1].split()[0]) == 0.25


def drive_problem(tmp_path, scale=2.0):
    """A gas at rest driven from +x by the pressure_input boundary with a pulse from a table file."""
    pb = sod_problem(ni=120, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk3", cfl=0.4, max_time=0.15)
    pb.cfg.bcs = ("pressure_input", "zero_gradient", "periodic", "periodic")
    pb.rho[...] = 1.0
    pb.p[...] = 0.1
    table = np.array([[0.0, 0.5, 0.25], [0.05, 0.5, 0.25], [0.06, 0.1, 0.25], [1.0, 0.1, 0.25]])
    path = os.path.join(str(tmp_path), "drive.dat")
    with open(path, "w") as fh:
        fh.write("# time [s]  pressure  density\n")
        fh.writelines(f"{float(t)!r} {float(p)!r}   {float(r)!r}\n" for t, p, r in table)
    return pb, path, (table[:, 0].copy(), (table[:, 1] * 1.0) * scale, (table[:, 2] * 1.0) * scale)


def test_host_reads_pressure_input_deck(tmp_path):
    """'pressure_input' at +x: the constant mode's constant_bc_pressure_value and bc_density reach the
    library's config (mod_input.f90:366-382); the boundary id is 4."""
    from fvleg_2d_amd.deck import write_deck

    pb, _, _ = drive_problem(tmp_path)
    pb.cfg.bc_pressure, pb.cfg.bc_density = 0.75, 0.125
    ini = write_deck(pb, str(tmp_path), "drive")
    r = subprocess.run([host_or_fail(), ini, "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    cfg_line = [l for l in r.stdout.splitlines() if l.startswith("config:")][0]
    assert cfg_line.split("=")[1].split()[3:7] == ["4", "1", "0", "0"]
    vals = [l for l in r.stdout.splitlines() if l.startswith("bc_pressure")][0].split("=")[1].split()
    assert [float(x) for x in vals] == [0.75, 0.125]


@pytest.mark.parametrize("extra, message", [
    ("[source_terms]\nenable_source_terms = true\n", "[source_terms] from the deck are not supported"),
])
def test_host_rejects_unsupported_deck_options(tmp_path, extra, message):
    """Deck options whose effect the host does not reproduce stop the run instead of being ignored."""
    from fvleg_2d_amd.deck import write_deck

    pb = sod_problem(ni=40, nj=4)
    ini = write_deck(pb, str(tmp_path), "sod")
    with open(ini, "a") as fh:
        fh.write("\n" + extra)
    r = subprocess.run([host_or_fail(), ini, "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode != 0
    assert message in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
def test_host_sod_matches_oracle(tmp_path, oracle, device_loop):
    """Config 1 (Sod 400x4, AUSM+, minmod, SSP-RK2, CFL 0.4, t = 0.2) through the Fortran host."""
    from fvleg_2d_amd.deck import read_state, write_deck

    pb = sod_problem(ni=400, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk2", cfl=0.4, max_time=0.2)
    ini = write_deck(pb, str(tmp_path), "sod")
    out = os.path.join(str(tmp_path), "final.h5")
    cmd = [host_or_fail(), ini, out] + (["--device-loop"] if device_loop else [])
    print("host:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_state(out)

    cfg = pb.cfg
    nh = cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    n, t, e = f.run(pb.max_time, 2 ** 31 - 1, pb.max_time)
    assert e == 0
    assert got["iteration"] == n and got["time"] == t
    for key, name in (("density", "rho"), ("rho_u", "rho_u"), ("rho_v", "rho_v"), ("rho_E", "rho_E")):
        assert np.array_equal(got[key], f.interior(name)), key


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
def test_host_curvilinear_deck_matches_oracle(tmp_path, oracle, device_loop):
    """A deck on a skewed (wavy) grid, HLLC + MLP5, SSP-RK3, to t = 0.03: the per-cell quad geometry of the
    deck's nodes runs through the host and matches the oracle bit for bit."""
    import dataclasses

    from fvleg_2d_amd.deck import read_state, write_deck
    from fvleg_2d_amd.problems import riemann2d, wavy_nodes

    pb = riemann2d(n=48, seed=4)
    nx, ny = wavy_nodes(48, 48, pb.cfg.n_ghost, 0.05)
    pb = dataclasses.replace(pb, node_x=nx, node_y=ny, max_time=0.03, contour_interval_dt=0.03)
    ini = write_deck(pb, str(tmp_path), "wavy")
    out = os.path.join(str(tmp_path), "final.h5")
    cmd = [host_or_fail(), ini, out] + (["--device-loop"] if device_loop else [])
    print("host:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_state(out)
    cfg = pb.cfg
    nh = cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    n, t, e = f.run(pb.max_time, 2 ** 31 - 1, pb.max_time)
    assert e == 0
    assert got["iteration"] == n and got["time"] == t
    for key, name in (("density", "rho"), ("rho_u", "rho_u"), ("rho_v", "rho_v"), ("rho_E", "rho_E")):
        assert np.array_equal(got[key], f.interior(name)), key


@pytest.mark.gpu
def test_host_writes_contours_on_the_reference_cadence(tmp_path):
    """main.f90:100-198 + contour_writer.f90:143-384: step_0000000.h5, one file per contour interval, and
    the last step; the last contour holds the final state with BC-filled ghost cells."""
    import glob

    from fvleg_2d_amd.deck import read_contour, read_state, write_deck

    pb = sod_problem(ni=100, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk2", cfl=0.4, max_time=0.2)
    pb.contour_interval_dt = 0.05
    ini = write_deck(pb, str(tmp_path), "sod")
    out = os.path.join(str(tmp_path), "final.h5")
    cdir = os.path.join(str(tmp_path), "results")
    r = subprocess.run([host_or_fail(), ini, out, "--contours", cdir, "--no-residual-hist"], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-2000:], flush=True)
    assert not os.path.exists(os.path.join(str(tmp_path), "residual_hist.csv"))
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted(glob.glob(os.path.join(cdir, "step_*.h5")))
    assert os.path.basename(files[0]) == "step_0000000.h5"
    assert len(files) >= 5, files
    c0 = read_contour(files[0])
    assert c0["iteration"] == 0 and c0["time"] == 0.0 and c0["n_ghost_layers"] == 2
    nh = 2
    assert np.array_equal(c0["density"][nh:-nh, nh:-nh], pb.rho)
    final = read_state(out)
    last = read_contour(files[-1])
    assert last["iteration"] == final["iteration"]
    assert np.array_equal(last["density"][nh:-nh, nh:-nh], final["density"])
    d = last["density"]
    assert np.array_equal(d[:nh, nh:-nh], np.repeat(d[nh:nh + 1, nh:-nh], nh, axis=0))  # zero-gradient -x
    assert np.array_equal(d[nh:-nh, :nh], d[nh:-nh, -2 * nh:-nh])  # periodic y
    times = [read_contour(f)["time"] for f in files]
    assert all(b > a for a, b in zip(times, times[1:]))


def oracle_main_loop(f, max_time, contour_dt, startup=None, initial=None):
    """main.f90:100-198 on an oracle Fluid: dt from the startup period (main.f90:141-145), the given first dt
    (main.f90:146-148) or the CFL, the contour-interval clip of the driver time; returns (iterations,
    driver time, residual rows)."""
    eps = np.finfo(float).eps
    rows, master, time, it, next_out = [], 0.0, 0.0, 0, 0.0 + contour_dt
    while time < max_time:
        if startup is not None:
            dt = startup[1] if time <= startup[0] else f.get_timestep()
        elif initial is not None and it == 0:
            dt = initial
        else:
            dt = f.get_timestep()
        assert f.integrate(dt) == 0
        master = master + dt
        it += 1
        rows.append([it, master, dt] + f.convergence())
        new_time = master
        if abs(time - next_out) < eps:
            next_out = next_out + contour_dt
        if new_time > next_out:
            new_time = time + abs(next_out - time)
        time = new_time
    return it, time, rows


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
@pytest.mark.parametrize("kind", ["startup", "initial"])
def test_host_startup_period_uses_constant_dt(tmp_path, oracle, kind, device_loop):
    """[time] enable_startup_time (main.f90:141-145): startup_delta_t while time <= startup_period, CFL after;
    [time] initial_delta_t (main.f90:146-148): the given dt on iteration 0, CFL after.  Host-driven loop and
    the device loop (cato_set_loop_dt) alike."""
    from fvleg_2d_amd.deck import read_state, write_deck

    pb = sod_problem(ni=100, nj=4, flux="AUSMPW+", limiter="van_leer", integrator="ssp_rk2", cfl=0.4, max_time=0.06)
    ini = write_deck(pb, str(tmp_path), "sod")
    period, dt0 = 0.01, 3e-4
    with open(ini, "a") as fh:
        if kind == "startup":
            fh.write(f"\n[time]\nenable_startup_time = true\nstartup_period = {period!r}\nstartup_delta_t = {dt0!r}\n")
        else:
            fh.write(f"\n[time]\ninitial_delta_t = {dt0!r}\n")
    out = os.path.join(str(tmp_path), "final.h5")
    cmd = [host_or_fail(), ini, out] + (["--device-loop"] if device_loop else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    got = read_state(out)
    hist = read_residual_hist(os.path.join(str(tmp_path), "residual_hist.csv"))

    cfg = pb.cfg
    nh = cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    if kind == "startup":
        n, t, rows = oracle_main_loop(f, pb.max_time, pb.max_time, startup=(period, dt0))
    else:
        n, t, rows = oracle_main_loop(f, pb.max_time, pb.max_time, initial=dt0)
    dts = np.array(rows)[:, 2]
    if kind == "startup":
        assert np.sum(dts == dt0) >= 30 and np.any(dts != dt0)  # both regimes ran
    else:
        assert dts[0] == dt0 and np.all(dts[1:] != dt0)
    assert got["iteration"] == n and got["time"] == t
    assert np.array_equal(hist[:, 2], [float(f"{x:.6e}") for x in dts])
    for key, name in (("density", "rho"), ("rho_u", "rho_u"), ("rho_v", "rho_v"), ("rho_E", "rho_E")):
        assert np.array_equal(got[key], f.interior(name)), key


def read_residual_hist(path):
    lines = open(path).read().splitlines()
    assert lines[0] == "iteration,time,dt,rho,rho_u,rho_v,rho_E"
    rows = []
    for line in lines[1:]:
        f = line.split(",")
        assert f[-1] == "" and len(f) == 8, line  # every value is followed by a comma (mod_fluid.f90:1217)
        assert all(len(x) == 16 for x in f[1:7]), line  # es16.6
        rows.append([float(x) for x in f[:7]])
    return np.array(rows)


@pytest.mark.gpu
def test_host_residual_history_matches_oracle(tmp_path, oracle):
    """residual_hist.csv (mod_fluid.f90:244-246,1207-1223) from the host-side and the device-side loop: the same
    text, one row per step, iteration/time/dt of the oracle's run and its convergence norms to es16.6 precision."""
    from fvleg_2d_amd.deck import write_deck

    pb = sod_problem(ni=200, nj=4, flux="SLAU", limiter="MLP3", integrator="ssp_rk3", cfl=0.4, max_time=0.05)
    ini = write_deck(pb, str(tmp_path), "sod")
    text = {}
    for mode in ("host", "device"):
        d = tmp_path / mode
        d.mkdir()
        cmd = [host_or_fail(), ini, str(d / "final.h5")] + (["--device-loop"] if mode == "device" else [])
        print("host:", " ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(d))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr
        text[mode] = open(d / "residual_hist.csv").read()
    assert text["host"] == text["device"]
    got = read_residual_hist(str(tmp_path / "host" / "residual_hist.csv"))

    cfg = pb.cfg
    nh = cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    want = np.array(oracle_main_loop(f, pb.max_time, pb.max_time)[2])
    assert got.shape == want.shape
    assert np.array_equal(got[:, 0], want[:, 0])
    assert np.allclose(got[:, 1:], want[:, 1:], rtol=6e-7, atol=0)


def run_host(args, cwd):
    r = subprocess.run([host_or_fail()] + args, capture_output=True, text=True, timeout=120, cwd=cwd)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    return r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
def test_host_restart_from_checkpoint_is_bit_exact(tmp_path, device_loop):
    """[restart] (main.f90:100-107, master_puppeteer.f90:106-121) from the conserved checkpoint the host writes:
    40 steps, then a restart to t_end, lands on exactly the state, time and iteration of the uninterrupted run.
    The device-side loop resumes from the restart time and iteration too (its driver clock starts at the
    fluid time); the residual history continues the iteration count."""
    from fvleg_2d_amd.deck import read_state, write_deck

    pb = sod_problem(ni=200, nj=4, flux="SLAU", limiter="minmod", integrator="ssp_rk3", cfl=0.4, max_time=0.1)
    d = str(tmp_path)
    full = write_deck(pb, d, "full")
    run_host([full, os.path.join(d, "a.h5")], d)
    part = write_deck(pb, d, "part", max_iterations=40)
    run_host([part, os.path.join(d, "b1.h5")], d)
    rest = write_deck(pb, d, "rest")
    with open(rest, "a") as fh:
        fh.write(f"\n[restart]\nrestart_from_file = true\nrestart_file = '{os.path.join(d, 'b1.h5')}'\n")
    out = run_host([rest, os.path.join(d, "b2.h5")] + (["--device-loop"] if device_loop else []), d)
    assert "Restart from" in out and "iteration 40" in out
    hist = read_residual_hist(os.path.join(d, "residual_hist.csv"))
    assert hist[0, 0] == 41 and hist[-1, 0] == read_state(os.path.join(d, "b2.h5"))["iteration"]
    a, b1, b2 = (read_state(os.path.join(d, n)) for n in ("a.h5", "b1.h5", "b2.h5"))
    assert b1["iteration"] == 40 and a["iteration"] > 40
    # same steps; the driver time differs by the reference's own rule: a restart sets next_output_time to
    # time + contour_interval_dt (main.f90:101), so the last step's clip to t_end (main.f90:188-194) is gone
    assert b2["iteration"] == a["iteration"]
    assert a["time"] == pb.max_time and b2["time"] > pb.max_time
    for key in ("density", "rho_u", "rho_v", "rho_E", "x_velocity", "y_velocity", "pressure"):
        assert np.array_equal(b2[key], a[key]), key


@pytest.mark.gpu
def test_host_restart_from_contour_matches_oracle(tmp_path, oracle):
    """A restart from a contour file (halo'd primitives, mod_fluid.f90:273-345): the run continues from that
    state, time and iteration like the oracle initialised from the same primitives."""
    import glob

    from fvleg_2d_amd.deck import read_contour, read_state, write_deck

    pb = sod_problem(ni=100, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk2", cfl=0.4, max_time=0.2)
    pb.contour_interval_dt = 0.05
    d = str(tmp_path)
    ini = write_deck(pb, d, "sod")
    run_host([ini, os.path.join(d, "a.h5"), "--contours", os.path.join(d, "results"), "--no-residual-hist"], d)
    files = sorted(glob.glob(os.path.join(d, "results", "step_*.h5")))
    mid = files[2]
    c = read_contour(mid)
    with open(ini, "a") as fh:
        fh.write(f"\n[restart]\nrestart_from_file = true\nrestart_file = '{mid}'\n")
    run_host([ini, os.path.join(d, "b.h5")], d)
    got = read_state(os.path.join(d, "b.h5"))

    cfg = pb.cfg
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.asfortranarray(c[k]) for k in ("density", "x_velocity", "y_velocity", "pressure")])
    eps = np.finfo(float).eps
    time = master = c["time"]
    it, next_out = c["iteration"], c["time"] + pb.contour_interval_dt
    while time < pb.max_time:  # main.f90:138-198 resumed at the contour's time (main.f90:100-104)
        dt = f.get_timestep()
        assert f.integrate(dt) == 0
        master = master + dt
        new_time = master
        if abs(time - next_out) < eps:
            next_out = next_out + pb.contour_interval_dt
        if new_time > next_out:
            new_time = time + abs(next_out - time)
        time = new_time
        it += 1
    assert got["iteration"] == it and got["time"] == time
    for key, name in (("density", "rho"), ("rho_u", "rho_u"), ("rho_v", "rho_v"), ("rho_E", "rho_E")):
        assert np.array_equal(got[key], f.interior(name)), key


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
def test_host_pressure_input_file_matches_oracle(tmp_path, oracle, device_loop):
    """A time-dependent pressure drive read from its file (header line, three columns, scale factor 2 on
    pressure and density, pressure_input_bc.f90:93-178) through the host, host- and device-side loops,
    against the oracle's main.f90 loop; within 1e-12 (the inlet rows' pow)."""
    from fvleg_2d_amd.deck import read_state, write_deck

    pb, path, table = drive_problem(tmp_path, scale=2.0)
    ini = write_deck(pb, str(tmp_path), "drive", pressure_input_file=path, pressure_scale_factor=2.0)
    out = os.path.join(str(tmp_path), "final.h5")
    cmd = [host_or_fail(), ini, out, "--no-residual-hist"] + (["--device-loop"] if device_loop else [])
    print("host:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_state(out)
    cfg = pb.cfg
    nh = cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl, pressure_table=table)
    f = oracle.Fluid(oc, oracle.Geometry(oc, pb.node_x, pb.node_y),
                     *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    n, t, e = f.run(pb.max_time, 2 ** 31 - 1, pb.max_time)
    assert e == 0 and got["iteration"] == n and abs(got["time"] - t) <= 1e-12 * t
    for key, name in (("density", "rho"), ("rho_u", "rho_u"), ("rho_v", "rho_v"), ("rho_E", "rho_E")):
        ref = f.interior(name)
        assert np.max(np.abs(got[key] - ref)) <= 1e-12 * np.max(np.abs(ref)), key
    assert np.max(f.interior("rho")) > 1.5  # the drive launched a shock into the gas


@pytest.mark.gpu
def test_host_subcycling_matches_oracle(tmp_path, oracle):
    """[time] allow_subcycle with a tolerance that the first steps of a Sod tube exceed: the host-side loop retries
    those steps with dt / 10 (mod_fluid.f90:816-823,856-887) and lands on the oracle's main.f90 loop bit for bit;
    the residual history shows the kept dt; --device-loop refuses the option."""
    from fvleg_2d_amd.deck import read_state, write_deck

    pb = sod_problem(ni=100, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk3", cfl=0.4, max_time=0.05)
    tol = 3e-3  # above the rho u norm of later steps (~2.5e-3), below the first ones' (~3.3e-3)
    ini = write_deck(pb, str(tmp_path), "sub")
    with open(ini, "a") as fh:
        fh.write(f"\n[time]\nallow_subcycle = true\nsubcycle_error_tolerance = {tol!r}\n")
    out = os.path.join(str(tmp_path), "final.h5")
    run_host([ini, out], str(tmp_path))
    got = read_state(out)
    hist = read_residual_hist(os.path.join(str(tmp_path), "residual_hist.csv"))
    cfg = pb.cfg
    nh = cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                       gamma=cfg.gamma, cfl=cfg.cfl, allow_subcycle=True, subcycle_tolerance=tol)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    n, t, e = f.run(pb.max_time, 2 ** 31 - 1, pb.max_time)
    assert e == 0 and got["iteration"] == n and got["time"] == t
    for key, name in (("density", "rho"), ("rho_u", "rho_u"), ("rho_v", "rho_v"), ("rho_E", "rho_E")):
        assert np.array_equal(got[key], f.interior(name)), key
    # the first step was retried: its kept dt is a tenth (or less) of the CFL dt of the next steps
    assert hist[0, 2] < 0.2 * hist[-1, 2] and np.all(hist[:, 3:] <= tol)
    r = subprocess.run([host_or_fail(), ini, out, "--device-loop"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode != 0 and "does not support allow_subcycle" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("units, plot_64bit", [("cgs", True), ("icf", False)])
def test_host_contour_units_and_precision(tmp_path, units, plot_64bit):
    """contour_writer_t%write_hdf5 (contour_writer.f90:193-384) in the deck's unit system (units.f90:94-143): icf
    writes ns, um, km/s and Mbar; every field carries description and units attributes; plot_64bit = false stores
    FP32 (write_2d_real_data, :544-587); /volume holds the halo'd cells' volumes.  A restart reads the icf contour
    back through its units attributes."""
    from fvleg_2d_amd import h5io
    from fvleg_2d_amd.deck import read_contour, read_state, write_deck

    pb = sod_problem(ni=60, nj=4, flux="AUSM+", limiter="minmod", integrator="ssp_rk2", cfl=0.4, max_time=0.05)
    d = str(tmp_path)
    ini = write_deck(pb, d, "u")
    with open(ini) as fh:
        text = fh.read()
    text = text.replace("[general]\n", f"[general]\nunits = '{units}'\n", 1)
    text += f"\n[io]\nplot_64bit = {'true' if plot_64bit else 'false'}\ncontour_interval_dt = 0.025\n"
    if units == "icf":
        text += "format = 'h5'\n"  # [io] format (mod_input.f90:429): no XDMF companion
    with open(ini, "w") as fh:
        fh.write(text)
    run_host([ini, os.path.join(d, "out.h5"), "--contours", os.path.join(d, "c")], d)
    out = read_state(os.path.join(d, "out.h5"))
    files = sorted(os.listdir(os.path.join(d, "c")))
    path = [f for f in files if f.endswith(".h5")][-1]
    c = read_contour(os.path.join(d, "c", path))
    xdmf = [f for f in files if f.endswith(".xdmf")]
    if units == "icf":
        assert not xdmf
    else:
        # contour_writer_t%write_xdmf: every DataItem names a dataset of its step file with matching dimensions
        import xml.etree.ElementTree as ET

        assert len(xdmf) == len([f for f in files if f.endswith(".h5")])
        root = ET.parse(os.path.join(d, "c", path.replace(".h5", ".xdmf"))).getroot()
        grid = root.find("Domain/Grid")
        assert grid.find("Time").get("Value").endswith(" s")
        names = [a.get("Name") for a in grid.findall("Attribute")]
        assert names == ["Density [g/cc]", "Sound Speed [cm/s]", "Pressure [barye]", "Velocity [cm/s]", "Mach",
                         "Volume [cc]"]
        with h5io.H5File(os.path.join(d, "c", path)) as f:
            for it in root.iter("DataItem"):
                if it.text and ":" in it.text:
                    fname, ds = it.text.strip().split(":")
                    assert fname == path
                    shape = tuple(int(k) for k in it.get("Dimensions").split())
                    assert f.read(ds).T.shape == shape, ds
    want = {"cgs": ("s", "cm", "cm/s", "barye"), "icf": ("ns", "um", "km/s", "Mbar")}[units]
    with h5io.H5File(os.path.join(d, "c", path)) as f:
        assert f.attr("/time", "units") == want[0] and f.attr("/time", "description") == "Simulation Time"
        assert f.attr("/delta_t", "units") == "seconds" and f.attr("/iteration", "units") == "dimensionless"
        assert f.attr("/x", "units") == want[1] and f.attr("/density", "units") == "g/cc"
        assert f.attr("/x_velocity", "units") == want[2] and f.attr("/sound_speed", "units") == want[2]
        assert f.attr("/pressure", "units") == want[3] and f.attr("/pressure", "description") == "Cell Pressure"
        assert f.itemsize("/pressure") == (8 if plot_64bit else 4) and f.itemsize("/volume") == (8 if plot_64bit else 4)
        vol = f.read("/volume")
    scale = {"cgs": (1.0, 1.0, 1.0, 1.0), "icf": (1e9, 1e4, 1e-5, 1e-12)}[units]
    nh = pb.cfg.n_ghost
    # the contour carries the master time (unclipped, master_puppeteer%time), the output the driver's
    assert 0.0 <= c["time"] / scale[0] - out["time"] <= c["delta_t"] * (1 + 1e-12)
    inner = (slice(nh, nh + pb.ni), slice(nh, nh + pb.nj))
    rtol = 1e-15 if plot_64bit else 1e-7
    assert np.allclose(c["pressure"][inner] / scale[3], out["pressure"], rtol=rtol, atol=0)
    assert np.allclose(c["x_velocity"][inner] / scale[2], out["x_velocity"], rtol=rtol, atol=1e-30)
    assert np.allclose(c["x"] / scale[1], pb.node_x, rtol=rtol, atol=1e-12)
    assert vol.shape == (pb.ni + 2 * nh, pb.nj + 2 * nh)
    assert np.allclose(vol, (1.0 / pb.ni) * (0.01 / pb.nj), rtol=1e-6)
    if units == "icf":
        rest = write_deck(pb, d, "r")
        with open(rest, "a") as fh:
            fh.write(f"\n[restart]\nrestart_from_file = true\nrestart_file = '{os.path.join(d, 'c', path)}'\n")
        out2 = run_host([rest, os.path.join(d, "r.h5")], d)
        assert "Restart from" in out2


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
def test_host_restart_at_max_iterations_takes_no_step(tmp_path, device_loop):
    """main.f90:138: a restart whose iteration is already max_iterations leaves the loop at once; the
    output repeats the restart state, time and iteration (the device loop is not entered either)."""
    from fvleg_2d_amd.deck import read_state, write_deck

    pb = sod_problem(ni=100, nj=4, flux="SLAU", limiter="minmod", integrator="ssp_rk3", cfl=0.4, max_time=0.1)
    d = str(tmp_path)
    part = write_deck(pb, d, "part", max_iterations=25)
    run_host([part, os.path.join(d, "b1.h5")], d)
    rest = write_deck(pb, d, "rest", max_iterations=25)
    with open(rest, "a") as fh:
        fh.write(f"\n[restart]\nrestart_from_file = true\nrestart_file = '{os.path.join(d, 'b1.h5')}'\n")
    run_host([rest, os.path.join(d, "b2.h5")] + (["--device-loop"] if device_loop else []), d)
    b1, b2 = read_state(os.path.join(d, "b1.h5")), read_state(os.path.join(d, "b2.h5"))
    assert b1["iteration"] == b2["iteration"] == 25
    assert b2["time"] == b1["time"]
    for key in ("density", "rho_u", "rho_v", "rho_E"):
        assert np.array_equal(b2[key], b1[key]), key


@pytest.mark.gpu
def test_host_writes_the_failing_state_to_a_contour(tmp_path):
    """main.f90:170-175: when integrate reports an error the state is written to a contour of the current
    iteration before the program stops with an error, also when that iteration already had one (here the
    first step: step_0000000.h5 then holds the failed state, not the initial one)."""
    import dataclasses

    from error_cases import ERROR_CASES, NI, NJ, error_case_fields
    from fvleg_2d_amd.deck import read_contour, write_deck
    from fvleg_2d_amd.problems import riemann2d

    case = ERROR_CASES["negative_density_and_pressure"]
    rho, u, v, p = error_case_fields("negative_density_and_pressure")
    base = riemann2d(n=NI, nj=NJ, flux="SLAU", limiter="minmod", integrator=case["integrator"])
    pb = dataclasses.replace(base, rho=rho, u=u, v=v, p=p, max_time=1.0, contour_interval_dt=1.0)
    ini = write_deck(pb, str(tmp_path), "fail")
    with open(ini, "a") as fh:
        fh.write(f"\n[time]\nuse_constant_delta_t = true\ninitial_delta_t = {case['dt']!r}\n")
    cdir = os.path.join(str(tmp_path), "results")
    r = subprocess.run([host_or_fail(), ini, os.path.join(str(tmp_path), "out.h5"), "--contours", cdir],
                       capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-1500:], flush=True)
    assert r.returncode != 0 and "error code 1" in r.stdout
    c = read_contour(os.path.join(cdir, "step_0000000.h5"))
    # contour scalars and fields are stored in single precision unless plot_64bit
    assert c["iteration"] == 0 and abs(c["time"] - case["dt"]) < 1e-7 and abs(c["delta_t"] - case["dt"]) < 1e-7
    nh = 2
    assert np.min(c["density"][nh:-nh, nh:-nh]) < 0.0
