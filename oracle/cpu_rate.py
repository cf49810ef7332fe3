"""CPU-baseline timing of the oracle (the bench's cpu_baseline legs; test infrastructure only).

    python -m oracle.cpu_rate --workload kh|sod|riemann [--n N] [--steps K]

runs the oracle's OpenMP build on a bounded sample of a workload and prints one JSON line: cell
updates per second, seconds, steps and the thread count (OMP_NUM_THREADS, set by the caller).  Each
leg runs in its own process so that its thread count is the one OpenMP starts with.
"""
import argparse
import json
import os
import sys
import time

os.environ["CATO_ORACLE_OMP"] = "1"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from fvleg_2d_amd.problems import kelvin_helmholtz, riemann2d, sod  # noqa: E402
from oracle import oracle as O  # noqa: E402


def problem(workload, n):
    if workload == "kh":
        return kelvin_helmholtz(n=n)
    if workload == "riemann":
        return riemann2d(n=n, seed=1)
    return sod()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["kh", "sod", "riemann"], default="kh")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=1, help="steps to time; 0 runs the workload to its end time")
    a = ap.parse_args()
    O.build()
    pb = problem(a.workload, a.n)
    cfg = pb.cfg
    oc = O.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                  gamma=cfg.gamma, cfl=cfg.cfl)
    og = O.Geometry(oc, pb.node_x, pb.node_y)
    nh = cfg.n_ghost
    f = O.Fluid(oc, og, *[np.asfortranarray(np.pad(x, nh)) for x in (pb.rho, pb.u, pb.v, pb.p)])
    t0 = time.perf_counter()
    if a.steps > 0:
        for _ in range(a.steps):
            f.integrate(f.get_timestep())
        steps = a.steps
    else:
        steps, _, err = f.run(pb.max_time, contour_interval_dt=pb.contour_interval_dt)
        assert err == 0
    el = time.perf_counter() - t0
    print(json.dumps({"workload": pb.name, "cells": pb.ni * pb.nj, "steps": int(steps), "seconds": el,
                      "cell_updates_per_s": pb.ni * pb.nj * steps / el,
                      "threads": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)}))


if __name__ == "__main__":
    main()
