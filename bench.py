"""Benchmark of the CATO hot path on one MI355X.

Headline workload (N=1): the Kelvin-Helmholtz shear layer on a 2048 x 2048
periodic grid, SLAU flux, MUSCL/TVD2 van Leer, SSP-RK3, CFL 0.5, FP64, 2000
steps.  A "step" is one SSP-RK3 step over the whole grid (three fused stage
launches + the CFL reduce); dt is the device-computed CFL step, so no host
synchronisation happens inside the timed region.  Inputs are resident in HBM
before timing.

The JSON line carries, beside the headline throughput:
  roofline        the fused stage kernel's algorithmic HBM bytes per launch
                  over its mean launch time (HIP events on the solve's stream)
  cpu_baseline    the C oracle on this host's cores (OpenMP over rows, a bounded sample)
  precision_form  the same workload with FP32 reconstruction/flux and FP64
                  state: throughput, roofline and its L1/Linf difference from
                  the FP64 state after the same steps
  shard_512x1024  the FP64 stage at one GPU's share of the headline grid split
                  4x2 over 8 GPUs (the shape a sharded solve would run per GPU)
  diagnostics     ms/step of the device loop with the per-step convergence
                  norms and history rows off and on
  riemann_1024    four-quadrant Riemann, 1024^2, HLLC + MLP5, SSP-RK3 to
                  t = 0.25, FP64 and FP32-flux forms, errors of the latter
  batched         256 independent Riemann solves (side log-uniform in
                  [64, 512]) in one persistent launch per stage vs one at a
                  time through the plain path
  served          Poisson arrivals of Riemann requests (side log-uniform in
                  [128, 2048], t_end in [0.1, 0.3]) through the serving loop
                  (FP32-flux) vs a one-at-a-time plain FP64 queue
`--forms none` keeps only the headline and its roofline/CPU legs.
`--mode copy` times a plain device copy of known size (the calibration run
for the rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
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
WARM_S = 1.0  # seconds of untimed stepping of the timed solve before its window opens


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_flux"])
    ap.add_argument("--tile", type=int, default=-1, help="-1: chosen by the autotuner")
    ap.add_argument("--forms", default="all", choices=["all", "none"])
    ap.add_argument("--mode", default="bench", choices=["bench", "copy"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--profile-only", action="store_true", help="run the timed loop only (for rocprofv3)")
    ap.add_argument("--served-requests", type=int, default=48, help="requests of the serving workload")
    return ap.parse_args()


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_leg(workload, n, steps, threads):
    """The C oracle (reference operation order, OpenMP over rows) on a bounded sample of a workload, in a
    process of its own started with OMP_NUM_THREADS = threads (oracle/cpu_rate.py)."""
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS=str(threads), CATO_ORACLE_OMP="1")
    out = subprocess.run([sys.executable, "-m", "oracle.cpu_rate", "--workload", workload, "--n", str(n),
                          "--steps", str(steps)], cwd=ROOT, env=env, check=True, capture_output=True, text=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def rel_errors(ref, got):
    """Per conserved variable: (mean |got - ref| / max |ref|, max |got - ref| / max |ref|)."""
    import numpy as np

    out = {}
    for name, r, g in zip(("rho", "rho_u", "rho_v", "rho_E"), ref, got):
        scale = max(float(np.max(np.abs(r))), 1e-300)
        d = np.abs(g - r)
        out[name] = {"l1": float(np.mean(d)) / scale, "linf": float(np.max(d)) / scale}
    return out


def timed_kh(n, precision, steps, warmup, tile, profile_only=False, nj=None, extra_steps=None):
    """Advance the KH workload; returns (wall seconds, stage-kernel mean ms, launches, final state).
    extra_steps: take exactly that many untimed steps after the warm-up instead of stepping for WARM_S
    seconds (a second form of the same workload ends on the same step count, so the states compare)."""
    from fvleg_2d_amd import Solver
    from fvleg_2d_amd.problems import kelvin_helmholtz

    pb = kelvin_helmholtz(n=n, precision=precision, nj=nj)
    tuned = None
    if tile < 0 and not profile_only:
        from fvleg_2d_amd.autotune import tune

        tile, times = tune(pb, steps=30, warmup=50)
        tuned = {"tile": tile, "stage_ms_by_tile": times}
    pb.cfg.tile = tile
    pb.meta["autotune"] = tuned
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
        s.advance(warmup)
        s.synchronize()
        # after the W warm-up steps, keep stepping (untimed) until WARM_S seconds have passed since the
        # solve started, so that the timed window does not open on clocks that idled during the
        # host-side set-up
        extra, t_w = 0, time.perf_counter()
        while not profile_only and (time.perf_counter() - t_w < WARM_S if extra_steps is None
                                    else extra < extra_steps):
            s.advance(8)
            s.synchronize()
            extra += 8
        pb.meta["extra_warmup_steps"] = extra
        # per-launch time of every fused stage in the timed region: HIP events on the
        # solve's stream (markers only; the host enqueues far ahead of the GPU)
        s.set_timing(not profile_only)
        t0 = time.perf_counter()
        s.advance(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        # the timed region must have taken all its steps: a step after an error (or past the
        # loop's end) returns at once and would inflate the rate
        it, _, err = s.loop_status()
        assert (it, err) == (steps, 0), f"timed region ran {it} of {steps} steps, error code {err}"
        if profile_only:
            return pb, wall, None, 0, None
        tm = s.timing()
        s.set_timing(False)
        assert tm["stage_launches"] == STAGES[pb.cfg.integrator] * steps, tm
        state = s.conserved()
    return pb, wall, tm["stage_ms"] / max(tm["stage_launches"], 1), tm["stage_launches"], state


def diagnostics_cost(pb, steps=500):
    """ms/step of the device-side loop (cato_run) on the same workload with the per-step convergence
    norms and history rows (calculate_convergence / write_residual_history) off and on."""
    from fvleg_2d_amd import Solver

    out = {"steps": steps, "loop": "cato_run, host poll every 32 steps"}
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        for on in (False, True, False, True):  # second pair after warm clocks; the last of each is kept
            s.set_diagnostics(on)
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            t0 = time.perf_counter()
            it, _, err = s.run(1e30, steps)
            wall = time.perf_counter() - t0
            assert (it, err) == (steps, 0)
            out["ms_per_step_on" if on else "ms_per_step_off"] = wall / steps * 1e3
            if on:
                assert len(s.history()) == steps
    out["overhead"] = out["ms_per_step_on"] / out["ms_per_step_off"] - 1.0
    return out


def device_graph_cost(pb_kh, kh_steps=200):
    """Device-side loop with the 32-step hipGraph replay against one launch per kernel: the Sod deck
    (400x4, AUSM+, minmod, SSP-RK2, to t = 0.2; launch-bound) through cato_run, and the KH workload through
    cato_advance.  Timing is off in both (per-launch events would force individual launches)."""
    from fvleg_2d_amd import Solver
    from fvleg_2d_amd.problems import sod

    out = {}
    pbs = sod()
    for label, pb in (("sod_400x4_run", pbs), (f"kh_{pb_kh.ni}x{pb_kh.nj}_advance", pb_kh)):
        res = {}
        with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
            for graph in (True, False, True, False):  # second pair after warm clocks; the last of each is kept
                s.set_device_graph(graph)
                s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
                s.synchronize()
                t0 = time.perf_counter()
                if pb is pbs:
                    n, _, err = s.run(pb.max_time, 0, pb.contour_interval_dt)
                    assert err == 0
                else:
                    n = kh_steps
                    s.advance(n)
                    s.synchronize()
                wall = time.perf_counter() - t0
                res["steps"] = n
                res["us_per_step_graph" if graph else "us_per_step_eager"] = wall / n * 1e6
        res["speedup"] = res["us_per_step_eager"] / res["us_per_step_graph"]
        out[label] = res
    return out


def roofline(pb, stage_avg_ms, launches, precision, n, measured_gbs=None):
    integ = pb.cfg.integrator
    cells = pb.ni * pb.nj
    bytes_per_launch = STEP_BYTES_PER_CELL[integ] / STAGES[integ] * cells
    achieved = bytes_per_launch / (stage_avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "stage_kernel_traffic.json")
    if os.path.exists(tpath):
        try:
            # calibrated HBM bytes per launch of the same workload, keyed by precision and tile
            traffic = json.load(open(tpath))[f"{n}_{precision}"][str(pb.cfg.tile)]["bytes_per_launch"]
        except (OSError, ValueError, KeyError):
            traffic = None
    return {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "kernel": "cato::stage_kernel (fused BC + reconstruction + flux + divergence + SSP stage)",
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "kernel_avg_ms": stage_avg_ms,
        "launches_timed": launches,
        "measured_ceiling_gbs": measured_gbs,
        "frac_of_measured_ceiling": (achieved / measured_gbs) if measured_gbs else None,
        "measured_ceiling_kernel": "cato_hbm_probe triad a = b + s c, 3 x 1 GiB FP64 arrays (median of 20)",
    }


def valu_figure(stage_avg_ms, precision, n, tile):
    """Vector-issue utilisation of the stage kernel: the VALU wave-instructions one launch issues (SQ_INSTS_VALU,
    profiles/stage_kernel_counters.json, counted per tile for the same workload) over this run's mean launch time,
    against 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.  None when no count matches."""
    cpath = os.path.join(ROOT, "profiles", "stage_kernel_counters.json")
    try:
        c = json.load(open(cpath)).get(f"{n}_{precision}", {}).get(str(tile))
    except Exception:
        c = None
    if not c:
        return None
    achieved = c["valu_insts_per_dispatch"] / (stage_avg_ms * 1e-3)
    peak = 1024 * 2.4e9 / 4
    return {"achieved": achieved, "peak": peak, "unit": "wave-instructions/s", "frac": achieved / peak,
            "insts_per_launch": c["valu_insts_per_dispatch"], "source": "profiles/stage_kernel_counters.json"}


def riemann_forms():
    """Config: 1024^2 four-quadrant Riemann, HLLC + MLP5, SSP-RK3, CFL 0.4, to t = 0.25."""
    from fvleg_2d_amd import Solver
    from fvleg_2d_amd.problems import riemann2d

    res = {}
    states = {}
    for prec in ("fp64", "fp32_flux"):
        pb = riemann2d(n=1024, seed=1, precision=prec)
        with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            s.synchronize()
            t0 = time.perf_counter()
            it, t, e = s.run(pb.max_time, 0, pb.max_time)
            wall = time.perf_counter() - t0
            states[prec] = s.conserved()
        res[prec] = {"steps": int(it), "t_end": t, "error_code": int(e), "wall_s": wall,
                     "cell_updates_per_s": pb.ni * pb.nj * it / wall}
    res["fp32_flux"]["error_vs_fp64"] = rel_errors(states["fp64"], states["fp32_flux"])
    res["note"] = "the FP64 path is bit-identical to the C oracle (GPU parity tests), so it is the FP64 reference here"
    return res


def hbm_probe(kind, gib=1, reps=20):
    """The library's streaming copy (kind 0) or triad (kind 1) over arrays of `gib` GiB: (GB/s, ms/launch)."""
    import ctypes as C

    from fvleg_2d_amd import _native as N

    gbs, ms = C.c_double(), C.c_double()
    N.check(N.load().cato_hbm_probe(int(kind), int(gib) << 30, int(reps), C.byref(gbs), C.byref(ms)),
            "cato_hbm_probe")
    return gbs.value, ms.value


def copy_calibration():
    """A plain device copy of known size (the library's streaming copy kernel, 1 GiB read and 1 GiB
    written per launch): FETCH_SIZE / WRITE_SIZE of this kernel calibrate the counters."""
    gbs, ms = hbm_probe(0, gib=1, reps=5)
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "launches": 5 + 3,
                      "copy_gb_per_s": gbs, "ms_per_launch": ms}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the solve does not "
              "shard across GPUs here.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy":
        copy_calibration()
        return

    say(f"[bench] KH {args.n}^2 {args.precision}: {args.warmup} warmup + {args.steps} timed steps")
    pb, wall, stage_ms, launches, state = timed_kh(args.n, args.precision, args.steps, args.warmup, args.tile,
                                                   args.profile_only)
    if args.profile_only:
        print(json.dumps({"profile_only": True, "steps": args.steps, "wall_s": wall}))
        return
    cells = pb.ni * pb.nj
    value = cells * args.steps / wall
    say("[bench] measured HBM ceiling (streaming triad, 3 x 1 GiB)")
    triad_gbs, _ = hbm_probe(1, gib=1, reps=20)
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
            "bytes_per_cell_step": STEP_BYTES_PER_CELL[pb.cfg.integrator],
            "tile": pb.cfg.tile,
            "autotune": pb.meta.get("autotune"),
        },
        "roofline": roofline(pb, stage_ms, launches, args.precision, args.n, triad_gbs),
        "valu": valu_figure(stage_ms, args.precision, args.n, pb.cfg.tile),
    }
    out["config"]["extra_warmup_steps"] = pb.meta.get("extra_warmup_steps")
    if args.forms == "all" and args.precision == "fp64":
        say("[bench] precision form (FP32 flux, FP64 state) on the same workload")
        pb32, wall32, ms32, l32, state32 = timed_kh(args.n, "fp32_flux", args.steps, args.warmup, args.tile,
                                                    extra_steps=pb.meta.get("extra_warmup_steps"))
        out["precision_form"] = {
            "value": cells * args.steps / wall32,
            "unit": "cell-updates/s",
            "ms_per_step": wall32 / args.steps * 1e3,
            "tile": pb32.cfg.tile,
            "autotune": pb32.meta.get("autotune"),
            "roofline": roofline(pb32, ms32, l32, "fp32_flux", args.n, triad_gbs),
            "valu": valu_figure(ms32, "fp32_flux", args.n, pb32.cfg.tile),
            "difference_from_fp64_state": rel_errors(state, state32),
            "difference_after_steps": args.warmup + pb32.meta.get("extra_warmup_steps", 0) + args.steps,
        }
        say("[bench] one GPU's share of the 2048^2 solve split over 8 GPUs: 512x1024")
        pbs, walls, mss, ls, _ = timed_kh(512, "fp64", 1000, args.warmup, pb.cfg.tile, nj=1024)
        out["shard_512x1024"] = {
            "value": pbs.ni * pbs.nj * 1000 / walls,
            "unit": "cell-updates/s",
            "steps": 1000,
            "ms_per_step": walls / 1000 * 1e3,
            "tile": pbs.cfg.tile,
            "stage_kernel_avg_ms": mss,
            "roofline_frac": (pbs.ni * pbs.nj * STEP_BYTES_PER_CELL[pbs.cfg.integrator] / 3) / (mss * 1e-3) / 8e12,
            "note": "the FP64 stage at the per-GPU shape of a 4x2 decomposition of the headline grid "
                    "(periodic on its own; no halo exchange)",
        }
        say("[bench] convergence diagnostics on/off (device loop)")
        out["diagnostics"] = diagnostics_cost(pb)
        say("[bench] device loop: 32-step hipGraph replay vs one launch per kernel")
        out["device_graph"] = device_graph_cost(pb)
        say("[bench] riemann 1024^2 HLLC/MLP5, FP64 and FP32-flux")
        out["riemann_1024"] = riemann_forms()
        from fvleg_2d_amd.benchforms import measure_batched, measure_served

        say("[bench] batched: 256 Riemann solves vs one at a time")
        out["batched"] = measure_batched(n_solves=256)
        say(f"[bench] served: {args.served_requests} Poisson arrivals vs a plain FP64 queue")
        out["served"] = measure_served(n=args.served_requests)
    if not args.no_cpu and os.environ.get("CATO_BENCH_NO_CPU") != "1":
        threads = host_threads()
        say(f"[bench] CPU baseline (C oracle, OpenMP over rows, {threads} threads)")
        leg = cpu_leg("kh", args.n, args.cpu_steps, threads)
        out["cpu_baseline"] = {
            "value": leg["cell_updates_per_s"],
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_steps} SSP-RK3 steps of the same {pb.ni}x{pb.nj} workload with the C oracle "
                      f"(reference operation order, -O3 -ffp-contract=off, OpenMP over rows on {threads} threads), "
                      f"{leg['seconds']:.1f} s",
        }
        if args.forms == "all":
            say("[bench] CPU legs: one core, Sod 400x4 full run, Riemann 1024^2 sample")
            legs = {
                f"kh_{args.n}_1core_1step": cpu_leg("kh", args.n, 1, 1),
                "sod_400x4_full_run_all_cores": cpu_leg("sod", 400, 0, threads),
                "sod_400x4_full_run_1core": cpu_leg("sod", 400, 0, 1),
                "riemann_1024_hllc_mlp5_2steps_all_cores": cpu_leg("riemann", 1024, 2, threads),
            }
            out["cpu_legs"] = legs
    print(json.dumps(out))


if __name__ == "__main__":
    main()
