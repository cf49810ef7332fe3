"""ctypes binding of libcato_hip.so (the C ABI declared in include/cato_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C fvleg_2d_amd/csrc``).  There is no fallback: if the library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcato_hip.so")

CATO_OK = 0

FLUX = {"AUSMPW+": 0, "SLAU": 1, "SLAU2": 2, "SD-SLAU": 3, "SD-SLAU2": 4,
        "Roe": 5, "S-Roe": 5, "SP-Roe": 5, "AUSM+": 6, "AUSM+-u": 6, "HLLC": 7, "M-AUSMPW+": 8}
LIMITER = {"minmod": 0, "superbee": 1, "van_leer": 2, "vanleer": 2, "MLP3": 3, "MLP5": 4}
BC = {"periodic": 0, "zero_gradient": 1, "symmetry": 2, "mirror": 2, "reflection": 2, "outflow": 3, "outlet": 3,
      "pressure_input": 4}
INTEGRATOR = {"ssp_rk2": 0, "ssp_rk3": 1, "ssp_rk33": 1, "rk(1,3)": 2, "rk(1,4)": 3, "rk(1,5)": 4,
              "rk(2,3)": 5, "rk(2,4)": 6, "rk(2,5)": 7}
PRECISION = {"fp64": 0, "fp32_flux": 1}


class CatoConfig(C.Structure):
    _fields_ = [
        ("flux", C.c_int), ("limiter", C.c_int), ("integrator", C.c_int), ("precision", C.c_int),
        ("bc_plus_x", C.c_int), ("bc_minus_x", C.c_int), ("bc_plus_y", C.c_int), ("bc_minus_y", C.c_int),
        ("n_ghost_layers", C.c_int), ("apply_low_mach_fix", C.c_int),
        ("gamma", C.c_double), ("cfl", C.c_double), ("tile", C.c_int), ("bc_pressure", C.c_double),
        ("bc_density", C.c_double),
    ]


class CatoGeom(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int),
                ("node_x", C.POINTER(C.c_double)), ("node_y", C.POINTER(C.c_double))]


_DP = C.POINTER(C.c_double)
_CTX = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "cato_last_error": (C.c_char_p, []),
    "cato_abi_version": (C.c_int, []),
    "cato_num_tiles": (C.c_int, []),
    "cato_tile_shape": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cato_ctx_create": (C.c_int, [C.POINTER(CatoConfig), C.POINTER(CatoGeom), C.POINTER(_CTX)]),
    "cato_ctx_destroy": (C.c_int, [_CTX]),
    "cato_set_primitive": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_set_conserved": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_get_conserved": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_set_energy_source": (C.c_int, [_CTX, _DP, C.c_int, C.c_int]),
    "cato_set_pressure_input": (C.c_int, [_CTX, _DP, _DP, _DP, C.c_int]),
    "cato_set_time": (C.c_int, [_CTX, C.c_double]),
    "cato_set_subcycle": (C.c_int, [_CTX, C.c_int, C.c_double]),
    "cato_last_dt": (C.c_int, [_CTX, _DP]),
    "cato_set_loop_dt": (C.c_int, [_CTX, C.c_double, C.c_int, C.c_double, C.c_double, C.c_longlong, C.c_int]),
    "cato_get_primitive": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_compute_dt": (C.c_int, [_CTX, _DP]),
    "cato_compute_dt_noalloc": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                          C.c_void_p]),
    "cato_step": (C.c_int, [_CTX, C.c_double, C.POINTER(C.c_int)]),
    "cato_run": (C.c_int, [_CTX, C.c_double, C.c_longlong, C.c_double, C.c_double, C.POINTER(C.c_longlong),
                           _DP, C.POINTER(C.c_int)]),
    "cato_advance": (C.c_int, [_CTX, C.c_longlong]),
    "cato_loop_status": (C.c_int, [_CTX, C.POINTER(C.c_longlong), _DP, C.POINTER(C.c_int)]),
    "cato_set_diagnostics": (C.c_int, [_CTX, C.c_int]),
    "cato_get_convergence": (C.c_int, [_CTX, _DP]),
    "cato_get_history": (C.c_int, [_CTX, _DP, C.c_longlong, C.POINTER(C.c_longlong)]),
    "cato_synchronize": (C.c_int, [_CTX]),
    "cato_set_timing": (C.c_int, [_CTX, C.c_int]),
    "cato_set_device_graph": (C.c_int, [_CTX, C.c_int]),
    "cato_get_timing": (C.c_int, [_CTX, _DP, C.POINTER(C.c_longlong), _DP, C.POINTER(C.c_longlong)]),
    "cato_apply_bc": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, C.c_int]),
    "cato_interpolate_edges": (C.c_int, [_CTX, _DP, C.c_int, _DP, _DP]),
    "cato_edge_fluxes": (C.c_int, [_CTX, _DP, _DP, C.c_int, _DP, _DP]),
    "cato_residual": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, C.c_int]),
    "cato_batch_create": (C.c_int, [C.c_int, C.POINTER(CatoConfig), C.POINTER(CatoGeom), C.c_int, C.c_int,
                                    C.POINTER(_CTX)]),
    "cato_batch_destroy": (C.c_int, [_CTX]),
    "cato_batch_set_primitive": (C.c_int, [_CTX, C.c_int, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_batch_get_conserved": (C.c_int, [_CTX, C.c_int, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_batch_set_pressure_input": (C.c_int, [_CTX, C.c_int, _DP, _DP, _DP, C.c_int]),
    "cato_batch_run": (C.c_int, [_CTX, _DP, _DP, C.c_longlong, C.c_int, C.POINTER(C.c_longlong), _DP,
                                 C.POINTER(C.c_int)]),
    "cato_device_bytes": (C.c_int, [C.POINTER(CatoConfig), C.POINTER(CatoGeom), C.POINTER(C.c_ulonglong)]),
    "cato_server_create": (C.c_int, [C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(_CTX)]),
    "cato_server_destroy": (C.c_int, [_CTX]),
    "cato_server_submit": (C.c_int, [_CTX, C.c_void_p, C.c_double]),
    "cato_server_tick": (C.c_int, [_CTX, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cato_server_poll": (C.c_int, [_CTX, C.c_int, C.POINTER(C.c_int), _DP, C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_int)]),
    "cato_server_clock": (C.c_double, [_CTX]),
    "cato_hbm_probe": (C.c_int, [C.c_int, C.c_ulonglong, C.c_int, _DP, _DP]),
    "cato_split_create": (C.c_int, [C.POINTER(CatoConfig), C.POINTER(CatoGeom), C.c_int, C.c_int,
                                    C.POINTER(_CTX)]),
    "cato_split_destroy": (C.c_int, [_CTX]),
    "cato_split_set_primitive": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_split_get_conserved": (C.c_int, [_CTX, _DP, _DP, _DP, _DP, C.c_int, C.c_int]),
    "cato_split_compute_dt": (C.c_int, [_CTX, _DP]),
    "cato_split_step": (C.c_int, [_CTX, C.c_double, C.POINTER(C.c_int)]),
    "cato_split_exchange_time": (C.c_int, [_CTX, _DP, C.POINTER(C.c_longlong)]),
}


class CatoRequest(C.Structure):
    _fields_ = [("cfg", CatoConfig), ("geom", CatoGeom),
                ("rho", _DP), ("u", _DP), ("v", _DP), ("p", _DP),
                ("max_time", C.c_double), ("contour_interval_dt", C.c_double),
                ("out_rho", _DP), ("out_rho_u", _DP), ("out_rho_v", _DP), ("out_rho_E", _DP)]


_lib = None


class CatoError(RuntimeError):
    pass


def load():
    """Load libcato_hip.so and declare every ABI signature.  Raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != CATO_OK:
        msg = load().cato_last_error().decode(errors="replace")
        raise CatoError(f"{what}: status {rc}: {msg}")
    return rc
