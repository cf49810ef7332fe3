"""Batched form: many independent solves advanced by one persistent launch per
RK stage (``cato_batch_*``), with steady-state steps replayed from a hipGraph."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .solver import SchemeConfig, _f64f, _p


class Batch:
    def __init__(self, problems, precision="fp64", tile=-1):
        """problems: sequence of objects with .cfg (SchemeConfig), .node_x, .node_y."""
        self.lib = N.load()
        self.n = len(problems)
        self.precision = precision
        self._keep = []
        cfgs = (N.CatoConfig * self.n)()
        geoms = (N.CatoGeom * self.n)()
        self.shapes = []
        for k, pb in enumerate(problems):
            cfgs[k] = pb.cfg.to_c()
            nx, ny = _f64f(pb.node_x), _f64f(pb.node_y)
            self._keep += [nx, ny]
            nh = pb.cfg.n_ghost
            ni, nj = nx.shape[0] - 1 - 2 * nh, nx.shape[1] - 1 - 2 * nh
            self.shapes.append((ni, nj))
            geoms[k] = N.CatoGeom(ni, nj, _p(nx), _p(ny))
        h = C.c_void_p()
        N.check(self.lib.cato_batch_create(self.n, cfgs, geoms, N.PRECISION[precision], int(tile), C.byref(h)),
                "cato_batch_create")
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.cato_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_primitive(self, k, rho, u, v, p):
        arrs = [_f64f(x) for x in (rho, u, v, p)]
        ni, nj = self.shapes[k]
        assert arrs[0].shape == (ni, nj)
        N.check(self.lib.cato_batch_set_primitive(self.h, k, *[_p(a) for a in arrs], ni, 0),
                "cato_batch_set_primitive")

    def set_pressure_input(self, k, table):
        """Time-tabled +x pressure input of solve k: table = (time, pressure, density) rows, already scaled;
        None returns to its config's constant bc_pressure / bc_density."""
        if table is None:
            N.check(self.lib.cato_batch_set_pressure_input(self.h, k, None, None, None, 0),
                    "cato_batch_set_pressure_input")
            return
        t, p, r = (np.ascontiguousarray(a, dtype=np.float64) for a in table)
        if not (len(t) == len(p) == len(r)):
            raise ValueError("time, pressure and density need the same length")
        N.check(self.lib.cato_batch_set_pressure_input(self.h, k, _p(t), _p(p), _p(r), len(t)),
                "cato_batch_set_pressure_input")

    def conserved(self, k):
        ni, nj = self.shapes[k]
        out = [np.zeros((ni, nj), order="F") for _ in range(4)]
        N.check(self.lib.cato_batch_get_conserved(self.h, k, *[_p(a) for a in out], ni, 0),
                "cato_batch_get_conserved")
        return out

    def run(self, max_time, contour_interval_dt=None, max_iterations=0, use_graph=True):
        mt = np.ascontiguousarray(np.broadcast_to(np.asarray(max_time, dtype=np.float64), (self.n,)))
        cd = np.ascontiguousarray(np.broadcast_to(
            np.asarray(contour_interval_dt if contour_interval_dt is not None else 0.0, dtype=np.float64),
            (self.n,)))
        it = np.zeros(self.n, dtype=np.int64)
        tt = np.zeros(self.n, dtype=np.float64)
        er = np.zeros(self.n, dtype=np.int32)
        N.check(self.lib.cato_batch_run(self.h, _p(mt), _p(cd), int(max_iterations), int(bool(use_graph)),
                                        it.ctypes.data_as(C.POINTER(C.c_longlong)), _p(tt),
                                        er.ctypes.data_as(C.POINTER(C.c_int))), "cato_batch_run")
        return it, tt, er
