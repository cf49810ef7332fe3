"""States that make one step of fluid_t%integrate fail, for the error-code checks (oracle and GPU).

Two streams moving apart (u = -V left of the middle, +V right of it) on a 20 x 10 zero-gradient block
with SLAU + minmod: with a dt well past the CFL limit the rarefaction empties the middle cells, and the
initial pressure and dt decide whether density, pressure or both go negative after the step or a stage
already produces a NaN.  A NaN may also be planted in the initial v.  Each case names the code of
error.f90:6-9 that the reference's check order gives (NaN at a stage's assign before negative density
before negative pressure); the oracle tests confirm it from the final state.
"""
import numpy as np

NI, NJ = 20, 10

ERROR_CASES = {
    "nan_planted_in_input": {"integrator": "ssp_rk2", "V": 0.0, "dt": 1e-3, "p0": 1.0, "nan": True, "code": 3},
    "nan_in_a_stage": {"integrator": "ssp_rk2", "V": 0.5, "dt": 0.1, "p0": 1.0, "code": 3},
    "negative_density_and_pressure": {"integrator": "ssp_rk2", "V": 0.5, "dt": 0.15, "p0": 1.0, "code": 1,
                                      "both": True},
    "negative_pressure_only": {"integrator": "ssp_rk2", "V": 0.5, "dt": 0.15, "p0": 0.1, "code": 2},
    "negative_density_only": {"integrator": "ssp_rk2", "V": 0.5, "dt": 0.15, "p0": 0.01, "code": 1},
    "nan_planted_and_negatives": {"integrator": "ssp_rk2", "V": 0.5, "dt": 0.15, "p0": 1.0, "nan": True,
                                  "code": 3},
    "negative_pressure_ssp_rk3": {"integrator": "ssp_rk3", "V": 4.0, "dt": 0.02, "p0": 1.0, "code": 2},
    "negative_density_rk24": {"integrator": "rk(2,4)", "V": 0.5, "dt": 0.2, "p0": 1.0, "code": 1},
    "negative_pressure_rk24": {"integrator": "rk(2,4)", "V": 1.0, "dt": 0.1, "p0": 1.0, "code": 2},
}


def error_case_fields(name):
    """Interior primitives (rho, u, v, p), each (NI, NJ) in Fortran order."""
    c = ERROR_CASES[name]
    rho = np.ones((NI, NJ), order="F")
    u = np.zeros((NI, NJ), order="F")
    u[: NI // 2, :] = -c["V"]
    u[NI // 2:, :] = c["V"]
    v = np.zeros((NI, NJ), order="F")
    if c.get("nan"):
        v[3, 2] = np.nan
    p = np.full((NI, NJ), c["p0"], order="F")
    return rho, u, v, p


def error_case_oracle(O, name):
    """(code, fluid) of one oracle step of the case."""
    c = ERROR_CASES[name]
    cfg = O.Config(NI, NJ, flux="SLAU", limiter="minmod", integrator=c["integrator"], bcs=("zero_gradient",) * 4)
    g = O.Geometry(cfg, *O.uniform_nodes(NI, NJ, cfg.nh))
    f = O.Fluid(cfg, g, *[np.asfortranarray(np.pad(a, cfg.nh)) for a in error_case_fields(name)])
    return f.integrate(c["dt"]), f
