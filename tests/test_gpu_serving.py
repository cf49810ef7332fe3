"""Serving form on the MI355X: continuous batching under an HBM budget.

Requests arrive while others are resident; each result must equal solving
that request alone through the plain single-solve path at the same precision
(bit for bit: the kernels and per-solve loop are shared), whatever it was
mixed with and whenever it was admitted.
"""
import sys

import numpy as np
import pytest

from fvleg_2d_amd import Solver
from fvleg_2d_amd.problems import riemann2d, sod
from fvleg_2d_amd.serve import Server, device_bytes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def requests(precision):
    mix = [("HLLC", "MLP5", 48, 0.05), ("AUSM+", "minmod", 40, 0.08), ("Roe", "MLP5", 56, 0.04),
           ("AUSMPW+", "minmod", 33, 0.06), ("SLAU2", "van_leer", 64, 0.03), ("SD-SLAU", "MLP3", 45, 0.05)]
    out = []
    for k, (fl, lim, n, t) in enumerate(mix):
        out.append(riemann2d(n=n, seed=500 + k, flux=fl, limiter=lim, max_time=t, precision=precision))
    return out


def alone(pb):
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        n, t, e = s.run(pb.max_time, 0, pb.max_time)
        return s.conserved(), n, e


def field_bytes(pb):
    return device_bytes(pb)


@pytest.mark.parametrize("precision", ["fp64", "fp32_flux"])
def test_result_same_alone_and_in_the_mix(precision):
    pbs = requests(precision)
    # room for two or three requests at a time: admission must wait for memory
    budget = 2.2 * max(field_bytes(pb) for pb in pbs)
    say("serving", precision, "budget", budget)
    with Server(budget, max_slots=8, precision=precision, steps_per_tick=4) as srv:
        ids = [srv.submit(pb) for pb in pbs[:3]]
        max_res = 0
        for _ in range(3):
            r, q, c = srv.tick()
            max_res = max(max_res, r)
        ids += [srv.submit(pb) for pb in pbs[3:]]  # arrive while others are on the GPU
        while True:
            r, q, c = srv.tick()
            max_res = max(max_res, r)
            if r == 0 and q == 0:
                break
        assert c == len(pbs)
        assert 2 <= max_res < len(pbs), max_res  # batched, and bounded by the budget
        got = [(srv.poll(i), srv.result(i)) for i in ids]
    for pb, ((done, tdone, n, e), res) in zip(pbs, got):
        ref, n0, e0 = alone(pb)
        say(pb.name, pb.cfg.flux, pb.cfg.limiter, "iters", n, n0, "err", e, e0, "done at", tdone)
        assert done and e == e0 == 0 and n == n0
        for v in range(4):
            assert np.array_equal(res[v], ref[v]), (pb.name, v)


def test_oversized_request_is_rejected_without_blocking():
    small = riemann2d(n=32, seed=3, flux="HLLC", limiter="minmod", max_time=0.02)
    big = riemann2d(n=96, seed=4, flux="HLLC", limiter="minmod", max_time=0.02)
    with Server(1.5 * field_bytes(small), max_slots=4) as srv:
        i_big = srv.submit(big)
        i_small = srv.submit(small)
        srv.drain()
        done, _, _, e_big = srv.poll(i_big)
        assert done and e_big == -1
        done, _, n, e = srv.poll(i_small)
        assert done and e == 0 and n > 0
        ref, n0, _ = alone(small)
        assert n == n0
        for v in range(4):
            assert np.array_equal(srv.result(i_small)[v], ref[v])


def test_sod_requests_share_the_server_with_2d_ones():
    a = sod(ni=200, nj=4, integrator="ssp_rk3", max_time=0.05)
    b = riemann2d(n=40, seed=9, flux="AUSM+", limiter="minmod", max_time=0.05)
    with Server(64e6, max_slots=4) as srv:
        ia, ib = srv.submit(a), srv.submit(b)
        srv.drain()
        for i, pb in ((ia, a), (ib, b)):
            done, _, n, e = srv.poll(i)
            ref, n0, _ = alone(pb)
            assert done and e == 0 and n == n0
            for v in range(4):
                assert np.array_equal(srv.result(i)[v], ref[v])


def test_integrator_mismatch_is_an_error():
    a = riemann2d(n=32, seed=1, flux="HLLC", limiter="minmod", integrator="ssp_rk3", max_time=0.01)
    b = riemann2d(n=32, seed=2, flux="HLLC", limiter="minmod", integrator="ssp_rk2", max_time=0.01)
    with Server(32e6, max_slots=2) as srv:
        srv.submit(a)
        with pytest.raises(RuntimeError):
            srv.submit(b)
        srv.drain()
