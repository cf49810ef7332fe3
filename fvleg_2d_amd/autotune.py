"""Tile autotuner of the fused stage kernel.

The stage kernel is compiled for every tile shape of ``cato_tile_shape``
(32x8, 32x16, 16x16, 64x8 cells per workgroup).  The best shape depends on the
scheme (LDS footprint and register pressure of the flux/limiter variant), the
precision form and the grid size (tile count against the 256 CUs), so it is
measured rather than predicted: each candidate advances a copy of the caller's
problem for a few steps and the mean stage-kernel time (HIP events) decides.
Choices are cached per (flux, limiter, integrator, precision, log2 cells) in
this process, and in the JSON file named by ``CATO_TILE_CACHE`` when set.

For the batched form the candidates are whole batches: a sample of the solves
is advanced for a few steps with each tile and the fastest wins.
"""
from __future__ import annotations

import copy
import ctypes as C
import json
import math
import os
import random
import time

from . import _native as N
from .solver import Solver

_cache: dict = {}


def num_tiles() -> int:
    return int(N.load().cato_num_tiles())


def tile_shape(t: int):
    tx, ty = C.c_int(), C.c_int()
    N.check(N.load().cato_tile_shape(int(t), C.byref(tx), C.byref(ty)), "cato_tile_shape")
    return tx.value, ty.value


def _key(cfg, ni, nj):
    return f"{cfg.flux}|{cfg.limiter}|{cfg.integrator}|{cfg.precision}|{int(round(math.log2(max(ni * nj, 1))))}"


def _load_file():
    path = os.environ.get("CATO_TILE_CACHE")
    if path and os.path.exists(path):
        try:
            _cache.update(json.load(open(path)))
        except (OSError, ValueError):
            pass


def _save_file():
    path = os.environ.get("CATO_TILE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)


def tune(pb, steps: int = 12, warmup: int = 3):
    """Time every tile on problem ``pb``; returns (best tile, {tile: mean stage ms})."""
    times = {}
    for t in range(num_tiles()):
        cfg = copy.copy(pb.cfg)
        cfg.tile = t
        with Solver(cfg, pb.node_x, pb.node_y) as s:
            s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
            s.advance(warmup)
            s.synchronize()
            s.set_timing(True)
            s.advance(steps)
            s.synchronize()
            tm = s.timing()
        times[t] = tm["stage_ms"] / max(tm["stage_launches"], 1)
    best = min(times, key=times.get)
    return best, times


def best_tile(pb, steps: int = 12) -> int:
    """Cached tile choice for problems like ``pb`` (same scheme, precision and size class)."""
    if not _cache:
        _load_file()
    k = _key(pb.cfg, pb.ni, pb.nj)
    if k not in _cache:
        best, times = tune(pb, steps=steps)
        _cache[k] = {"tile": best, "stage_ms": {str(t): v for t, v in times.items()}}
        _save_file()
    return int(_cache[k]["tile"])


def tune_batch(problems, sample: int = 48, steps: int = 12, precision: str = "fp64", seed: int = 0):
    """Time a sample of the batch with every tile; returns (best tile, {tile: seconds per step}).

    Each tile runs the sample for 4 and for 4 + steps steps from the same state; the difference is
    the cost of `steps` steps without the run's set-up (argument upload, first-dt reductions, the eager
    first step), which depends on the tile too and would otherwise decide a short trial."""
    from .batch import Batch

    pbs = list(problems)
    if len(pbs) > sample:
        pbs = random.Random(seed).sample(pbs, sample)
    times = {}
    for t in range(num_tiles()):
        with Batch(pbs, precision=precision, tile=t) as b:
            mt = [1e300] * len(pbs)
            for k, pb in enumerate(pbs):
                b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
            b.run(mt, mt, max_iterations=1)  # first step, graph capture
            wall = []
            for n in (4, 4 + steps):
                for k, pb in enumerate(pbs):
                    b.set_primitive(k, pb.rho, pb.u, pb.v, pb.p)
                t0 = time.perf_counter()
                b.run(mt, mt, max_iterations=n)
                wall.append(time.perf_counter() - t0)
            times[t] = (wall[1] - wall[0]) / steps
    best = min(times, key=times.get)
    return best, times
