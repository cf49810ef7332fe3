"""The reference's input deck for a problem: ``input.ini`` (the entries read by
src/lib/io/mod_input.f90:200-410) and the pycato HDF5 initial-condition file
(pycato/generate_initial_grids.py:603-672), as the Fortran host reads them.
"""
from __future__ import annotations

import os

import numpy as np

from . import h5io


def write_deck(pb, directory: str, name: str = "cato", max_iterations: int | None = None,
               pressure_input_file: str | None = None, pressure_scale_factor: float = 1.0) -> str:
    """Write ``<name>.h5`` and ``<name>.ini`` for problem ``pb`` into ``directory``; returns the ini path.

    Ghost cells of the deck are edge copies of the interior (the solver refills them through the
    boundary conditions before any use).  A +x ``pressure_input`` side reads ``pressure_input_file``
    when given, else the constant ``cfg.bc_pressure`` / ``cfg.bc_density`` (mod_input.f90:366-382)."""
    os.makedirs(directory, exist_ok=True)
    cfg = pb.cfg
    nh = cfg.n_ghost
    h5 = os.path.join(directory, name + ".h5")
    pad = [np.pad(np.asarray(a, dtype=np.float64), nh, mode="edge") for a in (pb.rho, pb.u, pb.v, pb.p)]
    h5io.write_initial_conditions(h5, nh, pb.node_x, pb.node_y, *pad)
    lines = [
        "[general]", f"title = '{pb.name}'", "",
        "[time]", f"max_time = {pb.max_time!r}", f"integration_strategy = '{cfg.integrator}'", f"cfl = {cfg.cfl!r}",
    ]
    if max_iterations:
        lines.append(f"max_iterations = {int(max_iterations)}")
    lines += [
        "", "[reference_state]", "non_dimensionalize = false", "",
        "[physics]", f"polytropic_index = {cfg.gamma!r}", "",
        "[scheme]", f"flux_solver = '{cfg.flux}'", "spatial_reconstruction = 'MUSCL'", f"limiter = '{cfg.limiter}'",
        f"apply_low_mach_fix = {'true' if cfg.apply_low_mach_fix else 'false'}", "",
        "[initial_conditions]", "read_from_file = true", f"initial_condition_file = '{h5}'", "",
        "[boundary_conditions]", f"plus_x = '{cfg.bcs[0]}'", f"minus_x = '{cfg.bcs[1]}'",
        f"plus_y = '{cfg.bcs[2]}'", f"minus_y = '{cfg.bcs[3]}'",
        *([f"outflow_ambient_pressure = {cfg.bc_pressure!r}"] if "outflow" in cfg.bcs else []),
        *_pressure_input_lines(cfg, pressure_input_file, pressure_scale_factor), "",
        "[io]", f"contour_interval_dt = {pb.contour_interval_dt or pb.max_time!r}", "",
    ]
    ini = os.path.join(directory, name + ".ini")
    with open(ini, "w") as f:
        f.write("\n".join(lines))
    return ini


def _pressure_input_lines(cfg, path, scale):
    if cfg.bcs[0] != "pressure_input":
        return []
    if path:
        return ["apply_constant_bc_pressure = false", f"bc_pressure_input_file = '{path}'",
                f"bc_pressure_scale_factor = {float(scale)!r}"]
    return ["apply_constant_bc_pressure = true", f"constant_bc_pressure_value = {cfg.bc_pressure!r}",
            f"bc_density = {cfg.bc_density!r}"]


def read_state(path: str) -> dict:
    """The Fortran host's output file: time, iteration and the interior fields."""
    with h5io.H5File(path, "r") as f:
        out = {k: f.read("/" + k) for k in ("time", "iteration", "density", "rho_u", "rho_v", "rho_E",
                                             "x_velocity", "y_velocity", "pressure")}
    out["time"] = float(np.asarray(out["time"]).reshape(-1)[0])
    out["iteration"] = int(np.asarray(out["iteration"]).reshape(-1)[0])
    return out


def read_contour(path: str) -> dict:
    """A contour file of the Fortran host (step_NNNNNNN.h5): scalars and halo'd fields."""
    out = {}
    with h5io.H5File(path, "r") as f:
        for k in ("iteration", "n_ghost_layers"):
            out[k] = int(f.read("/" + k, dtype=np.int32))
        for k in ("time", "delta_t"):
            out[k] = float(f.read("/" + k))
        for k in ("x", "y", "density", "x_velocity", "y_velocity", "pressure", "sound_speed"):
            out[k] = f.read("/" + k)
    return out
