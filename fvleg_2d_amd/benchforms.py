"""Throughput measurements of the batched and served forms (used by bench.py)."""
from __future__ import annotations

import time

import numpy as np

from .batch import Batch
from .problems import batch_riemann
from .solver import Solver


def plain_one_at_a_time(pbs):
    """Each solve through the plain single-solve API, strictly one after another."""
    t0 = time.perf_counter()
    results = []
    lat = []
    for pb in pbs:
        ts = time.perf_counter()
        with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            n, t, e = s.run(pb.max_time, 0, pb.max_time)
            results.append(s.conserved())
        lat.append(time.perf_counter() - ts)
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
    # warm the kernels of both paths on a tiny batch
    warm = batch_riemann(n_solves=4, seed=1, side_range=(32, 40), max_time=0.01)
    batched(warm, tile=tile)
    plain_one_at_a_time(warm[:1])
    b = batched(pbs, tile=tile)
    plain_s, plain_res, _ = plain_one_at_a_time(pbs)
    identical = None
    if check:
        identical = all(np.array_equal(a, r) for br, pr in zip(b["results"], plain_res) for a, r in zip(br, pr))
    cell_updates = float(sum(pb.ni * pb.nj * int(n) for pb, n in zip(pbs, b["iterations"])))
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
        "bit_identical_to_plain": identical,
        "errors": int(np.count_nonzero(b["errors"])),
    }
