"""Minimal HDF5 access through the HDF5 C library (ctypes), for the pycato deck
format that the reference reads (pycato/generate_initial_grids.py:603-672,
src/lib/grid/grid_block_2d.f90:194-265, src/lib/fluid/mod_fluid.f90:256-381):

  /n_ghost_layers  int scalar
  /x, /y           node coordinates, stored transposed (HDF5 dims (nj_nodes, ni_nodes))
  /density, /x_velocity, /y_velocity, /pressure   cell data incl. ghosts, transposed

Arrays come back as Fortran-ordered numpy arrays indexed (i, j).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import glob
import os

import numpy as np

_hid = C.c_int64
_lib = None


def _find_lib():
    cands = []
    env = os.environ.get("CATO_HDF5_LIB")
    if env:
        cands.append(env)
    name = ctypes.util.find_library("hdf5")
    if name:
        cands.append(name)
    for d in ("/opt/conda/lib", "/usr/lib/x86_64-linux-gnu", "/usr/lib64", "/usr/local/lib"):
        cands += sorted(glob.glob(os.path.join(d, "libhdf5.so*")))
        cands += sorted(glob.glob(os.path.join(d, "hdf5", "serial", "libhdf5.so*")))
    for c in cands:
        try:
            return C.CDLL(c)
        except OSError:
            continue
    raise OSError("HDF5 C library not found (set CATO_HDF5_LIB)")


def lib():
    global _lib
    if _lib is None:
        L = _find_lib()
        L.H5open.restype = C.c_int
        L.H5open()
        L.H5Fopen.restype = _hid
        L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, _hid]
        L.H5Fcreate.restype = _hid
        L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, _hid, _hid]
        L.H5Fclose.argtypes = [_hid]
        L.H5Dopen2.restype = _hid
        L.H5Dopen2.argtypes = [_hid, C.c_char_p, _hid]
        L.H5Dcreate2.restype = _hid
        L.H5Dcreate2.argtypes = [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]
        L.H5Dclose.argtypes = [_hid]
        L.H5Dget_space.restype = _hid
        L.H5Dget_space.argtypes = [_hid]
        L.H5Dread.argtypes = [_hid, _hid, _hid, _hid, _hid, C.c_void_p]
        L.H5Dwrite.argtypes = [_hid, _hid, _hid, _hid, _hid, C.c_void_p]
        L.H5Sget_simple_extent_ndims.argtypes = [_hid]
        L.H5Sget_simple_extent_dims.argtypes = [_hid, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.H5Screate_simple.restype = _hid
        L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.H5Screate.restype = _hid
        L.H5Screate.argtypes = [C.c_int]
        L.H5Sclose.argtypes = [_hid]
        L.H5Lexists.restype = C.c_int
        L.H5Lexists.argtypes = [_hid, C.c_char_p, _hid]
        L.H5Aopen_by_name.restype = _hid
        L.H5Aopen_by_name.argtypes = [_hid, C.c_char_p, C.c_char_p, _hid, _hid]
        L.H5Aget_type.restype = _hid
        L.H5Aget_type.argtypes = [_hid]
        L.H5Aread.argtypes = [_hid, _hid, C.c_void_p]
        L.H5Aclose.argtypes = [_hid]
        L.H5Dget_type.restype = _hid
        L.H5Dget_type.argtypes = [_hid]
        L.H5Tget_size.restype = C.c_size_t
        L.H5Tget_size.argtypes = [_hid]
        L.H5Tclose.argtypes = [_hid]
        _lib = L
    return _lib


def _native(name):
    return _hid.in_dll(lib(), name).value


class H5File:
    def __init__(self, path: str, mode: str = "r"):
        L = lib()
        self.path = path
        if mode == "r":
            self.h = L.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        else:
            self.h = L.H5Fcreate(path.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
        if self.h < 0:
            raise OSError(f"cannot open HDF5 file {path}")

    def close(self):
        if self.h >= 0:
            lib().H5Fclose(self.h)
            self.h = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exists(self, name: str) -> bool:
        return lib().H5Lexists(self.h, name.encode(), 0) > 0

    def read(self, name: str, dtype=np.float64) -> np.ndarray:
        """Dataset as a numpy array; 2D datasets come back transposed to (i, j) Fortran order."""
        L = lib()
        d = L.H5Dopen2(self.h, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        try:
            sp = L.H5Dget_space(d)
            nd = L.H5Sget_simple_extent_ndims(sp)
            dims = (C.c_uint64 * max(nd, 1))()
            L.H5Sget_simple_extent_dims(sp, dims, None)
            L.H5Sclose(sp)
            shape = tuple(int(dims[k]) for k in range(nd))
            mt = _native("H5T_NATIVE_DOUBLE_g") if dtype == np.float64 else _native("H5T_NATIVE_INT_g")
            out = np.empty(shape if nd else (), dtype=np.float64 if dtype == np.float64 else np.int32)
            if L.H5Dread(d, mt, 0, 0, 0, out.ctypes.data_as(C.c_void_p)) < 0:
                raise OSError(f"read of {name} failed")
        finally:
            L.H5Dclose(d)
        if out.ndim == 2:
            return np.asfortranarray(out.T)
        return out

    def attr(self, name: str, attr: str) -> str:
        """A fixed-length string attribute of a dataset."""
        L = lib()
        a = L.H5Aopen_by_name(self.h, name.encode(), attr.encode(), 0, 0)
        if a < 0:
            raise KeyError(f"{name}:{attr}")
        try:
            t = L.H5Aget_type(a)
            n = L.H5Tget_size(t)
            buf = C.create_string_buffer(n + 1)
            if L.H5Aread(a, t, buf) < 0:
                raise OSError(f"read of {name}:{attr} failed")
            L.H5Tclose(t)
        finally:
            L.H5Aclose(a)
        return buf.raw[:n].split(b"\0")[0].decode().rstrip()

    def itemsize(self, name: str) -> int:
        """Bytes per element as stored (4 for FP32 datasets, 8 for FP64)."""
        L = lib()
        d = L.H5Dopen2(self.h, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        try:
            t = L.H5Dget_type(d)
            n = int(L.H5Tget_size(t))
            L.H5Tclose(t)
        finally:
            L.H5Dclose(d)
        return n

    def write(self, name: str, data):
        """Write a scalar int or a 2D (i, j) float64 array (stored transposed, as pycato does)."""
        L = lib()
        arr = np.asarray(data)
        if arr.ndim == 0:
            sp = L.H5Screate(0)  # H5S_SCALAR
            mt = _native("H5T_NATIVE_INT_g")
            ft = _native("H5T_STD_I64LE_g")
            buf = np.array(int(arr), dtype=np.int32)
        else:
            buf = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).T)
            dims = (C.c_uint64 * buf.ndim)(*buf.shape)
            sp = L.H5Screate_simple(buf.ndim, dims, None)
            mt = _native("H5T_NATIVE_DOUBLE_g")
            ft = _native("H5T_IEEE_F64LE_g")
        d = L.H5Dcreate2(self.h, name.encode(), ft, sp, 0, 0, 0)
        if d < 0:
            L.H5Sclose(sp)
            raise OSError(f"cannot create dataset {name}")
        rc = L.H5Dwrite(d, mt, 0, 0, 0, buf.ctypes.data_as(C.c_void_p))
        L.H5Dclose(d)
        L.H5Sclose(sp)
        if rc < 0:
            raise OSError(f"write of {name} failed")


def read_initial_conditions(path: str) -> dict:
    """The datasets fluid_t/grid_block_2d_t read from an initial-condition deck."""
    with H5File(path) as f:
        nh = int(f.read("/n_ghost_layers", dtype=np.int32))
        out = {"n_ghost_layers": nh, "x": f.read("/x"), "y": f.read("/y")}
        for k in ("density", "x_velocity", "y_velocity", "pressure"):
            out[k] = f.read("/" + k)
    return out


def write_initial_conditions(path: str, n_ghost_layers: int, x, y, rho, u, v, p):
    """pycato's write_initial_hdf5 layout (units attributes omitted; cgs assumed)."""
    with H5File(path, "w") as f:
        f.write("/n_ghost_layers", n_ghost_layers)
        f.write("/x", x)
        f.write("/y", y)
        f.write("/density", rho)
        f.write("/x_velocity", u)
        f.write("/y_velocity", v)
        f.write("/pressure", p)
