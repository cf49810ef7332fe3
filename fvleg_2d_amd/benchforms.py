"""Throughput measurements of the batched and served forms (used by bench.py)."""
from __future__ import annotations

import copy
import dataclasses
import math
import sys
import time

import numpy as np

from .batch import Batch
from .problems import BATCH_FLUXES, BATCH_LIMITERS, batch_riemann, riemann2d
from .serve import Server
from .solver import Solver


def plain_one_at_a_time(pbs, parts=None):
    """Each solve through the plain single-solve API, strictly one after another.  parts (a dict), when
    given, collects the summed seconds of each solve's set-up (context, geometry, upload), run (the device
    loop: kernels, host polls, graph capture) and download (the result and the context's release)."""
    t0 = time.perf_counter()
    results = []
    lat = []
    for pb in pbs:
        ts = time.perf_counter()
        with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            tr = time.perf_counter()
            n, t, e = s.run(pb.max_time, 0, pb.max_time)
            td = time.perf_counter()
            results.append(s.conserved())
        te = time.perf_counter()
        if parts is not None:
            parts["setup_s"] = parts.get("setup_s", 0.0) + (tr - ts)
            parts["run_s"] = parts.get("run_s", 0.0) + (td - tr)
            parts["download_s"] = parts.get("download_s", 0.0) + (te - td)
        lat.append(te - ts)
    return time.perf_counter() - t0, results, lat


def batched(pbs, precision="fp64", tile=-1, use_graph=True):
    t0 = time.perf_counter()
    with Batch(pbs, precision=precision, tile=tile) as b:
        for k, pb in enumerate(pbs):
            b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
        t1 = time.perf_counter()
        mt = [pb.max_time for pb in pbs]
        it, tt, er = b.run(mt, mt, use_graph=use_graph)
        t2 = time.perf_counter()
        results = [b.conserved(k) for k in range(len(pbs))]
    t3 = time.perf_counter()
    return {"total_s": t3 - t0, "setup_s": t1 - t0, "run_s": t2 - t1, "iterations": it, "errors": er,
            "results": results}


def measure_batched(n_solves=256, seed=2024, side_range=(64, 512), max_time=0.2, tile=-1, check=True):
    pbs = batch_riemann(n_solves=n_solves, seed=seed, side_range=side_range, max_time=max_time)
    cells = int(sum(pb.ni * pb.nj for pb in pbs))
    tuned = None
    if tile < 0:
        from .autotune import tune_batch

        tile, times = tune_batch(pbs)
        tuned = {"tile": tile, "s_per_step_by_tile": times}
    # warm the kernels of both paths on a tiny batch
    warm = batch_riemann(n_solves=4, seed=1, side_range=(32, 40), max_time=0.01)
    batched(warm, tile=tile)
    plain_one_at_a_time(warm[:1])
    b = batched(pbs, tile=tile)
    parts = {}
    plain_s, plain_res, _ = plain_one_at_a_time(pbs, parts)
    identical = None
    if check:
        identical = all(np.array_equal(a, r) for br, pr in zip(b["results"], plain_res) for a, r in zip(br, pr))
    cell_updates = float(sum(pb.ni * pb.nj * int(n) for pb, n in zip(pbs, b["iterations"])))
    mix_rate, mix_rates = single_solve_mix_rate(pbs, b["iterations"], tile)
    return {
        "solves": n_solves,
        "cells_total": cells,
        "batched_solves_per_s": n_solves / b["total_s"],
        "batched_run_s": b["run_s"],
        "batched_total_s": b["total_s"],
        "plain_solves_per_s": n_solves / plain_s,
        "plain_total_s": plain_s,
        "speedup": plain_s / b["total_s"],
        "batched_cell_updates_per_s": cell_updates / b["run_s"],
        # the single-solve kernel on the same scheme mix at 1024^2 (weighted by the batch's own updates)
        "single_solve_cell_updates_per_s_same_mix_1024": mix_rate,
        "batched_over_single_solve_rate": cell_updates / b["run_s"] / mix_rate,
        "single_solve_rates_1024": mix_rates,
        "bit_identical_to_plain": identical,
        "errors": int(np.count_nonzero(b["errors"])),
        "tile": tile,
        "autotune": tuned,
        # where the plain path's time goes: per-solve set-up and download against the device loops; a
        # batch removes the first and most of the last, and runs the loops together
        "plain_breakdown_s": parts,
        "plain_fixed_cost_share": 1.0 - parts["run_s"] / plain_s,
        "speedup_ceiling_without_fixed_costs": plain_s / parts["run_s"],
    }


def single_solve_mix_rate(pbs, iterations, tile, side=1024, steps=60, warmup=10):
    """Cell-updates/s of the single-solve kernel on the batch's scheme mix at a large size: each flux x
    limiter pair of the batch timed on one side x side Riemann solve (FP64, the batch's tile), combined
    with the weights of the batch's own cell updates per pair (time = sum of updates / rate)."""
    weights = {}
    for pb, n in zip(pbs, iterations):
        key = (pb.cfg.flux, pb.cfg.limiter)
        weights[key] = weights.get(key, 0.0) + float(pb.ni * pb.nj * int(n))
    rates = {}
    for flux, lim in sorted(weights):
        pb = riemann2d(n=side, seed=3, flux=flux, limiter=lim)
        pb.cfg.tile = tile
        with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            s.advance(warmup)
            s.synchronize()
            t0 = time.perf_counter()
            s.advance(steps)
            s.synchronize()
            rates[f"{flux}/{lim}"] = side * side * steps / (time.perf_counter() - t0)
    total = sum(weights.values())
    t = sum(w / rates[f"{f}/{l}"] for (f, l), w in weights.items())
    return total / t, rates


def serving_requests(n=16, seed=4242, side_range=(128, 2048), t_range=(0.1, 0.3), precision="fp32_flux"):
    """Request mix of the serving workload: four-quadrant Riemann solves, side log-uniform in
    side_range, end time uniform in t_range, flux and reconstruction drawn as in the batched
    workload; plus unit-rate exponential inter-arrival gaps (scaled to the chosen rate later)."""
    rng = np.random.default_rng(seed)
    lo, hi = side_range
    reqs = []
    for k in range(n):
        side = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        flux = BATCH_FLUXES[int(rng.integers(len(BATCH_FLUXES)))]
        lim = BATCH_LIMITERS[int(rng.integers(len(BATCH_LIMITERS)))]
        t_end = float(rng.uniform(*t_range))
        pb = riemann2d(n=side, seed=int(rng.integers(1 << 30)), flux=flux, limiter=lim, max_time=t_end,
                       precision=precision)
        pb.name = f"req{k}_{side}_{flux}_{lim}"
        reqs.append(pb)
    gaps = rng.exponential(1.0, n)
    return reqs, gaps


def _as_precision(pb, precision):
    cfg = copy.copy(pb.cfg)
    cfg.precision = precision
    return dataclasses.replace(pb, cfg=cfg)


def progress(msg):
    """A progress line on stderr (long legs print one at least every few tens of seconds)."""
    print(msg, file=sys.stderr, flush=True)


def plain_queue_service_times(reqs):
    """Each request alone through the plain FP64 path (create, upload, run, download)."""
    times, results = [], []
    for k, pb in enumerate(reqs):
        if k and k % 16 == 0:
            progress(f"[served] plain FP64 queue: {k} of {len(reqs)} requests, {sum(times):.1f} s")
        q = _as_precision(pb, "fp64")
        t0 = time.perf_counter()
        with Solver(q.cfg, q.node_x, q.node_y) as s:
            s.set_primitive(q.rho, q.u, q.v, q.p)
            s.run(q.max_time, 0, q.max_time)
            res = s.conserved()
        times.append(time.perf_counter() - t0)
        results.append(res)
    return np.array(times), results


def fifo_tts(arrivals, service):
    """Time to solution of a single-server FIFO queue (start = max(arrival, previous end))."""
    end = 0.0
    tts = []
    for a, s in zip(arrivals, service):
        start = max(a, end)
        end = start + s
        tts.append(end - a)
    return np.array(tts), end


def served_run(reqs, arrivals, budget_bytes, max_slots=64, steps_per_tick=8, precision="fp32_flux"):
    """Drive the server in real time: request k is submitted once the clock passes arrivals[k]."""
    with Server(budget_bytes, max_slots=max_slots, precision=precision, steps_per_tick=steps_per_tick) as srv:
        t0 = srv.clock()
        ids = []
        nxt = 0
        n = len(reqs)
        ticks = 0
        max_res = 0
        reported = 0
        while True:
            now = srv.clock() - t0
            while nxt < n and arrivals[nxt] <= now:
                ids.append(srv.submit(reqs[nxt], arrival_time=t0 + arrivals[nxt]))
                nxt += 1
            r, q, c = srv.tick()
            ticks += 1
            max_res = max(max_res, r)
            if c == n:
                break
            if c >= reported + 16:
                reported = c - c % 16
                progress(f"[served] {c} of {n} requests done at {srv.clock() - t0:.1f} s")
            if r == 0 and q == 0 and nxt < n:
                time.sleep(max(0.0, min(0.01, arrivals[nxt] - (srv.clock() - t0))))
        polls = [srv.poll(i) for i in ids]
        results = [srv.result(i) for i in ids]
    done_t = np.array([p[1] - t0 for p in polls])
    tts = done_t - np.asarray(arrivals)
    errors = [p[3] for p in polls]
    return {"tts": tts, "makespan": float(done_t.max()), "results": results, "errors": errors, "ticks": ticks,
            "max_resident": max_res}


def rel_l1(ref, got):
    return max(float(np.mean(np.abs(g - r)) / max(np.max(np.abs(r)), 1e-300)) for r, g in zip(ref, got))


def measure_served(n=48, seed=4242, side_range=(128, 2048), load=0.9, budget_bytes=64e9, max_slots=64,
                   steps_per_tick=16):
    """Serving workload: Poisson arrivals at rate = load / (mean plain FP64 service time).

    Reports solves/s and p50/p99 time-to-solution of the server (FP32-flux/FP64-state) beside the
    one-at-a-time plain FP64 queue on the same arrival schedule, the saturated throughput of both
    (every request waiting at t=0), and the largest relative L1 difference of a served result
    from the plain FP64 result of the same request."""
    reqs, gaps = serving_requests(n=n, seed=seed, side_range=side_range)
    warm, _ = serving_requests(n=2, seed=1, side_range=(32, 40))
    served_run(warm, [0.0, 0.0], 1e9)
    plain_queue_service_times(warm[:1])
    service, plain_res = plain_queue_service_times(reqs)
    rate = load / float(service.mean())
    arrivals = np.cumsum(gaps) / rate
    arrivals -= arrivals[0]
    plain_tts, plain_end = fifo_tts(arrivals, service)
    sv = served_run(reqs, arrivals, budget_bytes, max_slots=max_slots, steps_per_tick=steps_per_tick)
    sat = served_run(reqs, np.zeros(n), budget_bytes, max_slots=max_slots, steps_per_tick=steps_per_tick)
    l1s = [rel_l1(p, s) for p, s in zip(plain_res, sv["results"])]
    l1 = max(l1s)
    pct = lambda a, q: float(np.percentile(a, q))
    return {
        "requests": n,
        "arrival_rate_per_s": rate,
        "offered_load_vs_plain": load,
        "hbm_budget_bytes": budget_bytes,
        "served_solves_per_s": n / sv["makespan"],
        "served_p50_tts_s": pct(sv["tts"], 50),
        "served_p99_tts_s": pct(sv["tts"], 99),
        "plain_fp64_solves_per_s": n / plain_end,
        "plain_fp64_p50_tts_s": pct(plain_tts, 50),
        "plain_fp64_p99_tts_s": pct(plain_tts, 99),
        "saturated_served_solves_per_s": n / sat["makespan"],
        "saturated_plain_fp64_solves_per_s": n / float(service.sum()),
        "saturated_speedup": float(service.sum()) / sat["makespan"],
        "p99_tts_ratio_plain_over_served": pct(plain_tts, 99) / pct(sv["tts"], 99),
        "max_rel_l1_vs_plain_fp64": l1,
        "rel_l1_vs_plain_fp64_per_request": [float(f"{x:.3g}") for x in l1s],
        "steps_per_tick": steps_per_tick,
        "percentile_note": f"p50 / p99 by linear interpolation over {n} requests",
        "max_resident": sv["max_resident"],
        "errors": int(sum(1 for e in sv["errors"] if e != 0)),
    }
