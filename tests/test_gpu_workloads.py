"""The benchmark workloads themselves on the MI355X, at their real sizes.

* Riemann 1024^2 (HLLC + MLP5, SSP-RK3, to t = 0.25) and Kelvin-Helmholtz 2048^2 (SLAU + van Leer,
  SSP-RK3, 2000 steps): the precision form (FP32 reconstruction and flux, FP64 state) against the FP64
  path, whose arithmetic is the oracle's bit for bit (the parity tests), with the L1 / L-infinity bounds
  stated below.
* Batched (256 Riemann solves, sides log-uniform in [64, 512], the {AUSM+, AUSMPW+, HLLC, Roe} x
  {minmod, MLP5} mix, to t = 0.2): a spread sample of 28 of the very 256 solves in one batch, each bit for
  bit the plain path's result, and the three smallest equal to the oracle's full run.
* Served (Poisson requests, sides log-uniform in [128, 2048], t_end in [0.1, 0.3], the same mix,
  FP32-flux): ten requests of the serving workload through the server under a budget that keeps a few
  resident, each bit for bit the plain precision-form solve alone and within the stated L1 of the FP64
  path; the smallest also served in FP64 and equal to the oracle.
"""
import sys

import numpy as np
import pytest

from fvleg_2d_amd import Solver
from fvleg_2d_amd.batch import Batch
from fvleg_2d_amd.benchforms import serving_requests
from fvleg_2d_amd.problems import batch_riemann, kelvin_helmholtz, riemann2d
from fvleg_2d_amd.serve import Server, device_bytes

pytestmark = pytest.mark.gpu

# precision form against FP64 after the full run, per conserved variable, relative to max |FP64|
PREC_L1 = 1e-4
PREC_LINF = 0.1
SERVED_L1 = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(got, ref):
    out = []
    for g, r in zip(got, ref):
        scale = max(float(np.max(np.abs(r))), 1e-300)
        d = np.abs(g - r)
        out.append((float(np.mean(d)) / scale, float(np.max(d)) / scale))
    return out


def run_plain(pb, precision=None, steps=None):
    cfg = pb.cfg
    if precision is not None:
        import copy

        cfg = copy.copy(pb.cfg)
        cfg.precision = precision
    with Solver(cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        if steps is None:
            n, t, e = s.run(pb.max_time, 0, pb.max_time)
        else:
            n, t, e = s.run(1e300, steps, 1e300)
        return s.conserved(), n, t, e


def oracle_run(O, pb):
    cfg = pb.cfg
    oc = O.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                  gamma=cfg.gamma, cfl=cfg.cfl)
    og = O.Geometry(oc, pb.node_x, pb.node_y)
    nh = cfg.n_ghost
    f = O.Fluid(oc, og, *[np.asfortranarray(np.pad(x, nh)) for x in (pb.rho, pb.u, pb.v, pb.p)])
    n, t, e = f.run(pb.max_time, contour_interval_dt=pb.max_time)
    return [f.interior(k) for k in ("rho", "rho_u", "rho_v", "rho_E")], n, t, e


@pytest.mark.parametrize("workload", ["riemann_1024_hllc_mlp5_t0.25", "kh_2048_slau_vanleer_2000_steps"])
def test_precision_form_error_at_full_size(workload):
    if workload.startswith("riemann"):
        pb, steps = riemann2d(n=1024, seed=1), None
    else:
        pb, steps = kelvin_helmholtz(n=2048), 2000
    say("precision form", workload)
    ref, n64, t64, e64 = run_plain(pb, "fp64", steps)
    got, n32, t32, e32 = run_plain(pb, "fp32_flux", steps)
    assert e64 == e32 == 0
    if steps is None:
        assert t64 == t32 == pb.max_time
    else:
        assert n64 == n32 == steps
    errs = rel_err(got, ref)
    say("  steps", n64, n32, "rel L1 / Linf per variable", errs)
    for l1, linf in errs:
        assert l1 <= PREC_L1 and linf <= PREC_LINF, errs


def c4_sample():
    """28 of the 256 solves of the batched workload, spread over its sizes (every 9th by side)."""
    pbs = batch_riemann(n_solves=256, seed=2024)
    order = sorted(range(len(pbs)), key=lambda k: (pbs[k].ni, k))
    pick = [order[k] for k in range(0, len(order), 9)][:28]
    return [pbs[k] for k in pick]


def test_batched_workload_sample_equals_plain_and_oracle(oracle):
    pbs = c4_sample()
    say("batched sample sides", [p.ni for p in pbs])
    assert min(p.ni for p in pbs) <= 66 and max(p.ni for p in pbs) >= 450
    with Batch(pbs) as b:
        for k, pb in enumerate(pbs):
            b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
        mt = [pb.max_time for pb in pbs]
        it, tt, er = b.run(mt, mt)
        got = [b.conserved(k) for k in range(len(pbs))]
    for k, pb in enumerate(pbs):
        ref, n, t, e = run_plain(pb)
        assert er[k] == e == 0 and it[k] == n and tt[k] == t == pb.max_time, (k, it[k], n, tt[k], t)
        for v in range(4):
            assert np.array_equal(got[k][v], ref[v]), (k, pb.name, v)
    for k in sorted(range(len(pbs)), key=lambda k: pbs[k].ni)[:3]:
        pb = pbs[k]
        say("  oracle", pb.name, pb.cfg.flux, pb.cfg.limiter)
        o, n, t, e = oracle_run(oracle, pb)
        assert e == 0 and n == it[k] and t == tt[k]
        for v in range(4):
            assert np.array_equal(got[k][v], o[v]), (pb.name, v)


C5_PICK = [8, 5, 15, 10, 14, 2, 1, 0, 12, 3]  # ten of the first sixteen requests: sides 142 ... 1256


def test_served_workload_requests(oracle):
    reqs, _ = serving_requests(n=16)
    pbs = [reqs[k] for k in C5_PICK]
    say("served requests", [(p.name, p.max_time) for p in pbs])
    # room for about three of the largest requests at once: admission waits for memory
    budget = 3.2 * max(device_bytes(pb) for pb in pbs)
    with Server(budget, max_slots=8, precision="fp32_flux", steps_per_tick=16) as srv:
        ids = [srv.submit(pb) for pb in pbs[:5]]
        for _ in range(4):
            srv.tick()
        ids += [srv.submit(pb) for pb in pbs[5:]]  # arrive while others are on the GPU
        srv.drain()
        got = [(srv.poll(i), srv.result(i)) for i in ids]
    for pb, ((done, _, n, e), res) in zip(pbs, got):
        alone, n32, _, e32 = run_plain(pb, "fp32_flux")
        ref64, n64, _, e64 = run_plain(pb, "fp64")
        assert done and e == e32 == e64 == 0 and n == n32
        for v in range(4):
            assert np.array_equal(res[v], alone[v]), (pb.name, v)
        errs = rel_err(res, ref64)
        say("  ", pb.name, "steps", n, n64, "rel L1 vs FP64", [round(a, 9) for a, _ in errs])
        assert all(l1 <= SERVED_L1 for l1, _ in errs), (pb.name, errs)
    # the smallest requests served in FP64: the oracle's full run, bit for bit
    small = sorted(pbs, key=lambda p: p.ni)[:3]
    import copy
    import dataclasses

    small64 = []
    for pb in small:
        cfg = copy.copy(pb.cfg)
        cfg.precision = "fp64"
        small64.append(dataclasses.replace(pb, cfg=cfg))
    with Server(4 * max(device_bytes(pb) for pb in small64), max_slots=4, precision="fp64",
                steps_per_tick=16) as srv:
        ids = [srv.submit(pb) for pb in small64]
        srv.drain()
        got = [(srv.poll(i), srv.result(i)) for i in ids]
    for pb, ((done, _, n, e), res) in zip(small64, got):
        say("  oracle", pb.name)
        o, no, to, eo = oracle_run(oracle, pb)
        assert done and e == eo == 0 and n == no
        for v in range(4):
            assert np.array_equal(res[v], o[v]), (pb.name, v)
