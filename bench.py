"""Benchmark of the CATO hot path on one MI355X.

Workload (N=1): the Kelvin-Helmholtz shear layer on a 2048 x 2048 periodic
grid, SLAU flux, MUSCL/TVD2 van Leer, SSP-RK3, CFL 0.5, FP64 (the
configuration the >= 70 % HBM-roofline target is quoted on).  A "step" is one
SSP-RK3 step over the whole grid (three fused stage launches + the CFL
reduce); dt is the device-computed CFL step, so no host synchronisation
happens inside the timed region.  Inputs are resident in HBM before timing.

Prints one JSON line (rank 0) with the throughput, the roofline of the fused
stage kernel (per-launch time from HIP events on the solve's stream) and a
CPU baseline (the C oracle on this host's cores, a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s & % HBM roofline; batched solves/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
# Compulsory HBM bytes per cell of a fused stage: stage 1 reads U^n and writes U^1
# (64 B); stages 2-3 read U^n and U^k and write U^k+1 (96 B).  SSP-RK3 step: 256 B.
STEP_BYTES_PER_CELL = {"ssp_rk3": (64 + 96 + 96), "ssp_rk2": (64 + 96)}
STAGES = {"ssp_rk3": 3, "ssp_rk2": 2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_flux"])
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--profile-only", action="store_true", help="run the timed loop only (for rocprofv3)")
    return ap.parse_args()


def cpu_baseline(pb, steps):
    """The C oracle (reference operation order, one thread) on the same workload."""
    import numpy as np

    from oracle import oracle as O

    O.build()
    cfg = pb.cfg
    oc = O.Config(pb.ni, pb.nj, flux=cfg.flux, limiter=cfg.limiter, integrator=cfg.integrator, bcs=cfg.bcs,
                  gamma=cfg.gamma, cfl=cfg.cfl)
    og = O.Geometry(oc, pb.node_x, pb.node_y)
    nh = cfg.n_ghost
    f = O.Fluid(oc, og, *[np.pad(a, nh) for a in (pb.rho, pb.u, pb.v, pb.p)])
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = f.get_timestep()
        f.integrate(dt)
    el = time.perf_counter() - t0
    return pb.ni * pb.nj * steps / el, el


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the solve does not "
              "shard across GPUs here.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from fvleg_2d_amd import Solver
    from fvleg_2d_amd.problems import kelvin_helmholtz

    pb = kelvin_helmholtz(n=args.n, precision=args.precision)
    pb.cfg.tile = args.tile
    cells = pb.ni * pb.nj
    s = Solver(pb.cfg, pb.node_x, pb.node_y)
    s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
    s.advance(args.warmup)
    s.synchronize()

    t0 = time.perf_counter()
    s.advance(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    if args.profile_only:
        print(json.dumps({"profile_only": True, "steps": args.steps, "wall_s": wall}))
        return

    # per-launch kernel time of the fused stage (HIP events on the solve's stream)
    tsteps = min(args.steps, 400)
    s.set_timing(True)
    s.advance(tsteps)
    s.synchronize()
    tm = s.timing()
    s.set_timing(False)
    s.close()

    integ = pb.cfg.integrator
    stage_avg_ms = tm["stage_ms"] / max(tm["stage_launches"], 1)
    bytes_per_launch = STEP_BYTES_PER_CELL[integ] / STAGES[integ] * cells
    achieved = bytes_per_launch / (stage_avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "stage_kernel_traffic.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath))
            if tj.get("n") == args.n and tj.get("precision") == args.precision:
                traffic = tj.get("bytes_per_launch")
        except Exception:
            traffic = None

    value = cells * args.steps / wall
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-flux/f64-state",
        "data": "synthetic (seeded Kelvin-Helmholtz initial condition)",
        "config": {
            "workload": f"kelvin_helmholtz_{pb.ni}x{pb.nj}_periodic_SLAU_vanleer_ssp_rk3_cfl0.5_{args.precision}",
            "cells": cells,
            "bytes_per_cell_step": STEP_BYTES_PER_CELL[integ],
            "tile": args.tile,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "cato::stage_kernel (fused BC+reconstruction+flux+divergence+SSP stage)",
            "kernel_avg_ms": stage_avg_ms,
            "launches_timed": tm["stage_launches"],
        },
    }
    if not args.no_cpu and os.environ.get("CATO_BENCH_NO_CPU") != "1":
        rate, el = cpu_baseline(pb, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": rate,
            "unit": "cell-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_steps} SSP-RK3 steps of the same {pb.ni}x{pb.nj} workload with the C oracle "
                      f"(one thread, -O2 -ffp-contract=off), {el:.1f} s",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
