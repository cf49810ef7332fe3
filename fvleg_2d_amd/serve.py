"""Serving form: a request loop with continuous batching (``cato_server_*``).

Requests are queued with :meth:`Server.submit`; every :meth:`Server.tick`
retires finished solves, admits queued requests while their device footprint
fits the HBM budget, and advances every resident solve by ``steps_per_tick``
steps in one persistent launch per RK stage.  Each result is the one the plain
single-solve path gives for the same request and precision.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .solver import _f64f, _p


def device_bytes(pb) -> int:
    """HBM footprint of one solve of problem ``pb`` (what the server charges against its budget)."""
    lib = N.load()
    nh = pb.cfg.n_ghost
    nx, ny = _f64f(pb.node_x), _f64f(pb.node_y)
    ni, nj = nx.shape[0] - 1 - 2 * nh, nx.shape[1] - 1 - 2 * nh
    cfg = pb.cfg.to_c()
    geom = N.CatoGeom(ni, nj, _p(nx), _p(ny))
    out = C.c_ulonglong()
    N.check(lib.cato_device_bytes(C.byref(cfg), C.byref(geom), C.byref(out)), "cato_device_bytes")
    return int(out.value)


class Server:
    def __init__(self, hbm_budget_bytes: float, max_slots: int = 64, precision: str = "fp64",
                 steps_per_tick: int = 8):
        self.lib = N.load()
        h = C.c_void_p()
        N.check(self.lib.cato_server_create(float(hbm_budget_bytes), int(max_slots), N.PRECISION[precision],
                                            int(steps_per_tick), C.byref(h)), "cato_server_create")
        self.h = h
        self.precision = precision
        self._req = {}  # id -> (keep-alive arrays, outputs)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cato_server_destroy(self.h)
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

    def clock(self) -> float:
        return float(self.lib.cato_server_clock(self.h))

    def submit(self, pb, max_time=None, arrival_time=None) -> int:
        """Queue problem ``pb`` (cfg, node_x, node_y, rho, u, v, p); returns the request id."""
        nh = pb.cfg.n_ghost
        nx, ny = _f64f(pb.node_x), _f64f(pb.node_y)
        ni, nj = nx.shape[0] - 1 - 2 * nh, nx.shape[1] - 1 - 2 * nh
        prim = [_f64f(a) for a in (pb.rho, pb.u, pb.v, pb.p)]
        for a in prim:
            if a.shape != (ni, nj):
                raise ValueError(f"primitive shape {a.shape} != {(ni, nj)}")
        out = [np.zeros((ni, nj), order="F") for _ in range(4)]
        mt = float(pb.max_time if max_time is None else max_time)
        req = N.CatoRequest()
        req.cfg = pb.cfg.to_c()
        req.geom = N.CatoGeom(ni, nj, _p(nx), _p(ny))
        req.rho, req.u, req.v, req.p = (_p(a) for a in prim)
        req.max_time = mt
        req.contour_interval_dt = mt
        req.out_rho, req.out_rho_u, req.out_rho_v, req.out_rho_E = (_p(a) for a in out)
        t = self.clock() if arrival_time is None else float(arrival_time)
        rid = self.lib.cato_server_submit(self.h, C.byref(req), t)
        if rid < 0:
            N.check(rid, "cato_server_submit")
        self._req[rid] = out  # inputs are copied by the library; outputs must live until completion
        return rid

    def tick(self):
        """One tick; returns (resident, queued, completed_total)."""
        r, q, c = C.c_int(), C.c_int(), C.c_int()
        N.check(self.lib.cato_server_tick(self.h, C.byref(r), C.byref(q), C.byref(c)), "cato_server_tick")
        return r.value, q.value, c.value

    def poll(self, rid: int):
        """(done, completion_time, iterations, error_code) of request ``rid``."""
        d, t, n, e = C.c_int(), C.c_double(), C.c_longlong(), C.c_int()
        N.check(self.lib.cato_server_poll(self.h, int(rid), C.byref(d), C.byref(t), C.byref(n), C.byref(e)),
                "cato_server_poll")
        return bool(d.value), t.value, n.value, e.value

    def result(self, rid: int):
        """Conserved interior fields (rho, rho u, rho v, rho E) of a completed request."""
        return self._req[rid]

    def drain(self, max_ticks: int = 10 ** 9):
        """Tick until nothing is resident or queued; returns the number of ticks."""
        for k in range(max_ticks):
            r, q, _ = self.tick()
            if r == 0 and q == 0:
                return k + 1
        raise RuntimeError("server did not drain")
