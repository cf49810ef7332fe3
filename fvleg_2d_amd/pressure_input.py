"""The pressure input file of the +x ``pressure_input`` boundary
(pressure_input_bc_t%read_pressure_input, src/lib/boundary_conditions/pressure_input_bc.f90:93-178).

A first line starting with '#' is a header; every following non-blank line holds time [s],
pressure [barye] and density [g/cc].  Pressure and density are multiplied by
``bc_pressure_scale_factor`` (pressure_input_bc.f90:163-164; time, pressure and density are
in code units, the reference state this library supports having unit scales).  The times must
increase strictly (linear_interp_1d%initialize, finterp's check_inputs).
"""
from __future__ import annotations

import re

import numpy as np


def parse_pressure_input(text: str, scale_factor: float = 1.0):
    """(time, pressure, density) arrays of a pressure input file's text."""
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if lines and lines[0].lstrip().startswith("#"):
        lines = lines[1:]
    rows = []
    for n, ln in enumerate(lines, 1):
        vals = [v for v in re.split(r"[,\s]+", ln.strip()) if v]
        if len(vals) < 3:
            raise ValueError(f"pressure input row {n}: expected time, pressure and density, got {ln!r}")
        # list-directed input: each row is one read of three values (the rest of the line is skipped)
        rows.append([float(v.replace("d", "e").replace("D", "e")) for v in vals[:3]])
    if len(rows) < 2:
        raise ValueError("cannot use linear interpolation for only one point")
    a = np.array(rows, dtype=np.float64)
    t = a[:, 0].copy()
    if not np.all(np.diff(t) > 0.0):
        raise ValueError("pressure input times are not strictly increasing")
    return t, (a[:, 1] * 1.0) * float(scale_factor), (a[:, 2] * 1.0) * float(scale_factor)


def read_pressure_input(path: str, scale_factor: float = 1.0):
    with open(path) as f:
        return parse_pressure_input(f.read(), scale_factor)
