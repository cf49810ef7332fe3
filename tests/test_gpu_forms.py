"""Batched and reduced-precision forms on the MI355X.

- A batch of heterogeneous solves (every flux, both reconstruction families,
  uneven sizes and end times) gives, per solve, exactly the result of solving
  it alone through the plain API, with and without hipGraph replay.
- The FP32-flux / FP64-state form stays close to the FP64 oracle; its L1 and
  L-infinity errors per conserved variable are printed for the record.
"""
import sys

import numpy as np
import pytest

from fvleg_2d_amd import Solver
from fvleg_2d_amd.batch import Batch
from fvleg_2d_amd.problems import batch_riemann, kelvin_helmholtz, riemann2d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve_alone(pb, max_time):
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        n, t, e = s.run(max_time, 0, max_time)
        return s.conserved(), n, t, e


def mixed_batch():
    pbs = batch_riemann(n_solves=10, seed=77, side_range=(20, 70), max_time=0.05)
    fl = ["AUSMPW+", "SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2", "Roe", "AUSM+", "HLLC", "M-AUSMPW+", "Roe"]
    lims = ["minmod", "superbee", "van_leer", "MLP3", "MLP5", "minmod", "MLP5", "MLP3", "van_leer", "superbee"]
    out = []
    for k, pb in enumerate(pbs):
        q = riemann2d(n=pb.ni, seed=100 + k, flux=fl[k], limiter=lims[k], max_time=0.03 + 0.01 * (k % 4))
        out.append(q)
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_solving_alone(use_graph):
    pbs = mixed_batch()
    say("batch", [(p.ni, p.cfg.flux, p.cfg.limiter) for p in pbs], "graph", use_graph)
    with Batch(pbs) as b:
        for k, pb in enumerate(pbs):
            b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
        mt = [pb.max_time for pb in pbs]
        it, tt, er = b.run(mt, mt, use_graph=use_graph)
        got = [b.conserved(k) for k in range(len(pbs))]
    for k, pb in enumerate(pbs):
        ref, n, t, e = solve_alone(pb, pb.max_time)
        assert er[k] == e == 0
        assert it[k] == n and tt[k] == t, (k, it[k], n, tt[k], t)
        for v in range(4):
            assert np.array_equal(got[k][v], ref[v]), (k, v)


def test_batch_matches_oracle_for_one_member(oracle):
    pbs = mixed_batch()[:3]
    with Batch(pbs) as b:
        for k, pb in enumerate(pbs):
            b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
        mt = [pb.max_time for pb in pbs]
        b.run(mt, mt)
        got = b.conserved(1)
    pb = pbs[1]
    nh = pb.cfg.n_ghost
    oc = oracle.Config(pb.ni, pb.nj, flux=pb.cfg.flux, limiter=pb.cfg.limiter, integrator=pb.cfg.integrator,
                       bcs=pb.cfg.bcs, gamma=pb.cfg.gamma, cfl=pb.cfg.cfl)
    og = oracle.Geometry(oc, pb.node_x, pb.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    f.run(pb.max_time, 2 ** 31 - 1, pb.max_time)
    for v, n in enumerate(("rho", "rho_u", "rho_v", "rho_E")):
        assert np.array_equal(got[v], f.interior(n))


def errors_vs(ref, got):
    out = []
    for r, g in zip(ref, got):
        scale = max(np.max(np.abs(r)), 1e-300)
        out.append((float(np.mean(np.abs(g - r)) / scale), float(np.max(np.abs(g - r)) / scale)))
    return out


def test_fp32_flux_form_riemann(oracle):
    """Four-quadrant Riemann (HLLC, MLP5) at 96^2 to t = 0.25: FP32 reconstruction+flux vs the FP64 oracle."""
    pb64 = riemann2d(n=96, seed=1)
    pb32 = riemann2d(n=96, seed=1, precision="fp32_flux")
    nh = pb64.cfg.n_ghost
    oc = oracle.Config(96, 96, flux="HLLC", limiter="MLP5", integrator="ssp_rk3", bcs=pb64.cfg.bcs, cfl=0.4)
    og = oracle.Geometry(oc, pb64.node_x, pb64.node_y)
    f = oracle.Fluid(oc, og, *[np.pad(a, nh) for a in (pb64.rho, pb64.u, pb64.v, pb64.p)])
    f.run(0.25, 2 ** 31 - 1, 0.25)
    ref = [f.interior(n) for n in ("rho", "rho_u", "rho_v", "rho_E")]
    got, n, t, e = solve_alone(pb32, 0.25)
    assert e == 0
    errs = errors_vs(ref, got)
    say("fp32-flux riemann96 rel L1/Linf per variable:", errs)
    for l1, linf in errs:
        assert l1 < 2e-4 and linf < 5e-2


def test_fp32_flux_form_kh_conservation():
    """Periodic KH at 128^2, 100 steps: the FP64 state keeps mass and momentum totals to round-off."""
    pb = kelvin_helmholtz(n=128, steps=100, precision="fp32_flux")
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        u0 = s.conserved()
        s.run(1e300, 100, 1e300)
        u1 = s.conserved()
    for k in range(3):
        tot0, tot1 = u0[k].sum(), u1[k].sum()
        mag = np.abs(u0[k]).sum()
        say("fp32-flux KH total", k, tot0, tot1)
        assert abs(tot1 - tot0) <= 1e-10 * mag


@pytest.mark.parametrize("precision", ["fp64", "fp32_flux"])
def test_results_independent_of_tile_and_autotune(precision):
    """Every tile shape computes each cell with the same arithmetic: identical states after 20 steps;
    the autotuner returns one of them with a positive time for each."""
    import copy

    from fvleg_2d_amd.autotune import num_tiles, tune

    pb = kelvin_helmholtz(n=96, steps=20, precision=precision)
    states = []
    for t in range(num_tiles()):
        cfg = copy.copy(pb.cfg)
        cfg.tile = t
        with Solver(cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            s.advance(20)
            states.append(s.conserved())
    for st in states[1:]:
        for v in range(4):
            assert np.array_equal(st[v], states[0][v])
    best, times = tune(pb, steps=4, warmup=1)
    say("autotune", precision, best, times)
    assert best in times and all(v > 0 for v in times.values())


FLUXES = ["AUSMPW+", "SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2", "Roe", "AUSM+", "HLLC", "M-AUSMPW+"]


@pytest.mark.parametrize("precision", ["fp64", "fp32_flux"])
@pytest.mark.parametrize("flux", FLUXES)
def test_every_flux_independent_of_tile(precision, flux):
    """Every flux and both reconstruction families give identical states on every tile shape.  FP64 tiles
    of one cell per thread run most edges in rounds of one direction (the flux's normal folded at compile
    time) and the tile-boundary edges with the normal read at run time, the other tiles all edges at run
    time; the precision form runs one loop over all edges on every tile.  70 x 53 cells (partial tiles),
    zero-gradient boundaries, 12 SSP-RK3 steps."""
    import copy

    from fvleg_2d_amd.autotune import num_tiles

    for lim in ("van_leer", "MLP5"):
        pb = riemann2d(n=70, nj=53, seed=5, flux=flux, limiter=lim, precision=precision)
        states = []
        for t in range(num_tiles()):
            cfg = copy.copy(pb.cfg)
            cfg.tile = t
            with Solver(cfg, pb.node_x, pb.node_y) as s:
                s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
                s.advance(12)
                states.append(s.conserved())
        say("tiles", precision, flux, lim, "max |diff|",
            max(float(np.abs(st[v] - states[0][v]).max()) for st in states[1:] for v in range(4)))
        for st in states[1:]:
            for v in range(4):
                assert np.array_equal(st[v], states[0][v]), (flux, lim, v)


def test_batch_results_independent_of_tile():
    from fvleg_2d_amd.autotune import num_tiles

    pbs = mixed_batch()[:5]
    got = []
    for t in range(num_tiles()):
        with Batch(pbs, tile=t) as b:
            for k, pb in enumerate(pbs):
                b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
            mt = [pb.max_time for pb in pbs]
            b.run(mt, mt)
            got.append([b.conserved(k) for k in range(len(pbs))])
    for g in got[1:]:
        for k in range(len(pbs)):
            for v in range(4):
                assert np.array_equal(g[k][v], got[0][k][v])
