"""The C ABI library loads and exports every entry point include/cato_hip.h declares.
No compute call is made here (CPU container, no HIP device)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cato_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cato_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("cato_ctx_create", "cato_step", "cato_run", "cato_residual", "cato_apply_bc",
              "cato_interpolate_edges", "cato_edge_fluxes", "cato_compute_dt"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from fvleg_2d_amd import _native

    lib = _native.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    from fvleg_2d_amd import _native

    assert set(declared_symbols()) <= set(_native.SIGNATURES), set(declared_symbols()) - set(_native.SIGNATURES)


def test_abi_version_and_tiles_without_device():
    from fvleg_2d_amd import _native

    lib = _native.load()
    assert lib.cato_abi_version() == 4
    n = lib.cato_num_tiles()
    assert n >= 1
    tx, ty = ctypes.c_int(), ctypes.c_int()
    for t in range(n):
        assert lib.cato_tile_shape(t, ctypes.byref(tx), ctypes.byref(ty)) == 0
        assert tx.value * ty.value >= 128
    assert lib.cato_tile_shape(n, ctypes.byref(tx), ctypes.byref(ty)) != 0


def test_struct_layout_matches_header():
    """cato_config / cato_geom field order and sizes as declared in the header."""
    from fvleg_2d_amd import _native as N

    text = open(HEADER).read()
    body = re.search(r"typedef struct cato_config \{(.*?)\} cato_config;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ty, rest = decl.split(None, 1)
        names += [n.strip() for n in rest.split(",")]
    assert names == [f[0] for f in N.CatoConfig._fields_]


def test_struct_offsets_match_c_compiler(tmp_path):
    """sizeof/offsetof of the ABI structs as gcc lays them out == the ctypes mirrors."""
    import subprocess

    from fvleg_2d_amd import _native as N

    structs = {"cato_config": N.CatoConfig, "cato_geom": N.CatoGeom, "cato_request": N.CatoRequest}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "cato_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.split("\n") if line)
    for cname, py in structs.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert int(out[f"{cname}.{f[0]}"]) == getattr(py, f[0]).offset, (cname, f[0])


def test_device_bytes_needs_no_device():
    """The HBM footprint query validates the configuration on the host only."""
    from fvleg_2d_amd.problems import riemann2d
    from fvleg_2d_amd.serve import device_bytes

    small = riemann2d(n=32, seed=1, flux="HLLC", limiter="minmod")
    big = riemann2d(n=64, seed=1, flux="HLLC", limiter="MLP5")
    b_small, b_big = device_bytes(small), device_bytes(big)
    assert b_small >= 24 * 8 * 32 * 32  # three conserved + three primitive buffer sets, FP64
    assert b_big > 3 * b_small
