"""cato_compute_dt_noalloc (the allocation-free CFL entry point) against cato_compute_dt.

The caller owns every buffer the entry point reads or writes: this test allocates the device
arrays itself through the HIP runtime the library uses, uploads the primitives that
cato_compute_dt reduces (cato_get_primitive, the reference's halo'd layout) and compares the dt
the allocation-free path writes into device memory with cato_compute_dt's, bit for bit, over calls
in a row: a grid large enough that the 4096 partial minima of the fixed scratch are each shared by
several workgroups' worth of cells, then smaller grids (the scratch keeps the previous call's
contents, which must not leak into the next result).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from fvleg_2d_amd import Solver
from fvleg_2d_amd import _native as N
from fvleg_2d_amd.problems import kelvin_helmholtz, riemann2d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceArray:
    """A device copy of a host float64 array, allocated by the caller (this test) with hipMalloc."""

    hip = None

    def __init__(self, host: np.ndarray):
        if DeviceArray.hip is None:
            DeviceArray.hip = C.CDLL("libamdhip64.so.7")
        self.n = host.size
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * max(self.n, 1))) == 0
        src = np.asfortranarray(host, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, src.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0

    def read(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


def noalloc_dt(s: Solver, prims, dt_dev: DeviceArray, halo: int) -> float:
    """cato_compute_dt_noalloc on device copies of the interior prims, stored with `halo` ghost layers of
    NaN around them (the layout of the reference's field_2d_t%data when halo = n_ghost_layers): the
    ghosts must not be read."""
    lib = N.load()
    ld = s.ni + 2 * halo
    dev = [DeviceArray(np.pad(a, halo, constant_values=np.nan)) for a in prims]
    try:
        N.check(lib.cato_compute_dt_noalloc(s.h, dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, ld, halo, dt_dev.ptr),
                "cato_compute_dt_noalloc")
        s.synchronize()
        return float(dt_dev.read()[0])
    finally:
        for d in dev:
            d.free()


CASES = [
    # (label, problem): the first has 1100 x 1100 cells, 4727 workgroups of 256 cells > the 4096 partials
    ("riemann_1100x1000_hllc_mlp5", lambda: riemann2d(n=1100, seed=3, flux="HLLC", limiter="MLP5")),
    ("kh_96x64_fp64", lambda: kelvin_helmholtz(n=96, nj=64)),
    ("kh_40x24_fp32_flux", lambda: kelvin_helmholtz(n=40, nj=24, precision="fp32_flux")),
]


def test_noalloc_dt_equals_compute_dt_on_calls_in_a_row():
    dt_dev = DeviceArray(np.full(1, np.nan))
    try:
        for label, make in CASES:
            say("noalloc dt", label)
            pb = make()
            with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
                s.set_primitive(pb.rho, pb.u, pb.v, pb.p)
                s.advance(3)  # a state that the stages produced (Mach-filtered primitives)
                s.synchronize()
                dt_ref = s.compute_dt()
                prims = s.primitive()[:4]
                for halo in (s.nh, 0):
                    got = noalloc_dt(s, prims, dt_dev, halo)
                    say("  dt", halo, got, dt_ref)
                    assert got == dt_ref, (label, halo, got, dt_ref)
    finally:
        dt_dev.free()


def test_noalloc_rejects_a_bad_layout():
    pb = kelvin_helmholtz(n=16, nj=16)
    with Solver(pb.cfg, pb.node_x, pb.node_y) as s:
        lib = N.load()
        x = DeviceArray(np.zeros(4))
        try:
            rc = lib.cato_compute_dt_noalloc(s.h, x.ptr, x.ptr, x.ptr, x.ptr, 10, 2, x.ptr)  # ld < ni + 2 * halo
            assert rc == -1
            rc = lib.cato_compute_dt_noalloc(s.h, x.ptr, None, x.ptr, x.ptr, 20, 2, x.ptr)
            assert rc == -1
        finally:
            x.free()
