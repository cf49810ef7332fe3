"""Pins of the oracle's edge fluxes and error codes that the mathematics (or the reference's own
stencil definitions) fix, beyond consistency and symmetry.

Each test evaluates the oracle at single edges with hand-chosen left/right states and checks a
property that a plausible transcription mistake (a dropped term, a wrong sign or index, a swapped
split function, a textbook variant in place of the reference's) would break:

* supersonic upwinding: where every split function of a scheme saturates, the flux is the upwind
  physical flux (AUSM+, AUSMPW+, M-AUSMPW+, HLLC) or the closed form the SLAU family reduces to;
* HLLC (Toro 2009, ch. 10; the reference has no usable HLLC) resolves isolated contacts and shear
  waves exactly and agrees with Toro's alternative star-flux form (eq. 10.41-10.42) on general
  subsonic states: a different arrangement of the same algebra;
* the pressure-split polynomials with the reference's constants: alpha = 1/8 in AUSMPW+
  (ausmpw_plus_solver.fypp:61, 486-505), none in M-AUSMPW+ (m_ausmpw_plus_solver.fypp:655-662),
  alpha = 3/16 and beta = 1/8 in AUSM+ (ausm_plus_solver.f90:89-128), extracted from the flux;
* the reference's Roe variant (roe_solver.fypp:475-483): F_L, F_R use the Roe-averaged normal
  velocity, which by the Roe property leaves a closed-form offset from F(U_L) in supersonic flow,
  and the pressure-sensor correction (495-511) as an isolated momentum term;
* the cells each shock sensor reads (slau_solver.fypp:410-452, ausmpw_plus_solver.fypp:432-441,
  m_ausmpw_plus_solver.fypp:672-681, roe_solver.fypp:495-507);
* the error codes of fluid_t%integrate in the reference's check order (error.f90:6-9,
  mod_fluid.f90:718-737, field_2d.f90:825-867).
"""
import math

import numpy as np
import pytest

from error_cases import ERROR_CASES, error_case_oracle

G = 1.4
NI = NJ = 6
I0, J0 = 3, 3  # the edge under test: i-edge between cells (3,3) and (4,3); j-edge between (3,3) and (3,4)


def _block(O, flux, low_mach_fix=True):
    cfg = O.Config(NI, NJ, flux=flux, limiter="minmod", gamma=G, apply_low_mach_fix=low_mach_fix)
    geom = O.Geometry(cfg, *O.uniform_nodes(NI, NJ, cfg.nh))
    return cfg, geom


def edge_flux(O, flux, L, R, direction, pcell=None, sb=None, low_mach_fix=True):
    """Flux of `flux` at one edge whose reconstructed states are L and R = (rho, u, v, p).
    pcell: halo'd cell pressures read by the shock sensors (uniform 1 when None); sb: the superbee
    states of M-AUSMPW+ (the same states when None)."""
    cfg, geom = _block(O, flux, low_mach_fix)
    edges = []
    for k in range(4):
        e = np.zeros((2, NI + 1, NJ + 1), order="F")
        e[0], e[1] = L[k], R[k]
        edges += [e, e.copy(order="F")]
    if flux == "M-AUSMPW+":
        S = (L, R) if sb is None else sb
        for k in range(4):
            e = np.zeros((2, NI + 1, NJ + 1), order="F")
            e[0], e[1] = S[0][k], S[1][k]
            edges += [e, e.copy(order="F")]
    p = cfg.field()
    p[:] = 1.0
    if pcell is not None:
        p[:] = pcell
    fi, fj = O.edge_fluxes(cfg, geom, p, p, p, p, edges)
    return fi[:, I0, J0 - 1].copy() if direction == 0 else fj[:, I0 - 1, J0].copy()


def normal(direction):
    return (1.0, 0.0) if direction == 0 else (0.0, 1.0)


def phys(S, direction):
    rho, u, v, p = S
    nx, ny = normal(direction)
    E = p / (G - 1.0) + 0.5 * rho * (u * u + v * v)
    q = u * nx + v * ny
    return np.array([rho * q, rho * u * q + p * nx, rho * v * q + p * ny, q * (E + p)])


def close(a, b, rtol):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) <= rtol * scale


def along(direction, un, ut, rho, p):
    """A state with normal velocity un and tangential velocity ut for the edge's direction."""
    return (rho, un, ut, p) if direction == 0 else (rho, ut, un, p)


# ---------------------------------------------------------------------------------------------
# supersonic upwinding
# ---------------------------------------------------------------------------------------------
SUPERSONIC = [
    # (L, R) in (rho, normal velocity, tangential velocity, p), both sides at normal Mach > 2
    ((1.0, 3.2, 0.3, 1.0), (0.8, 3.6, -0.4, 0.7)),
    ((0.5, 5.0, -1.0, 0.4), (1.3, 4.1, 0.9, 1.6)),
]


@pytest.mark.parametrize("flux", ["AUSM+", "AUSMPW+", "M-AUSMPW+", "HLLC"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("case", range(len(SUPERSONIC)))
@pytest.mark.parametrize("sense", [1, -1])
def test_supersonic_flow_takes_the_upwind_physical_flux(oracle, flux, direction, case, sense):
    """|M| > 1 on both sides: every split function saturates (P+ = 1, P- = 0, M+ = M, M- = 0 for flow
    in +n) and the flux is F(U_L); flow in -n gives F(U_R)."""
    a, b = SUPERSONIC[case]
    L = along(direction, sense * a[1], a[2], a[0], a[3])
    R = along(direction, sense * b[1], b[2], b[0], b[3])
    F = edge_flux(oracle, flux, L, R, direction)
    up = L if sense > 0 else R
    assert close(F, phys(up, direction), 1e-14), (F, phys(up, direction))


@pytest.mark.parametrize("flux", ["SLAU", "SLAU2", "SD-SLAU", "SD-SLAU2"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("case", range(len(SUPERSONIC)))
@pytest.mark.parametrize("sense", [1, -1])
def test_slau_family_supersonic_closed_form(oracle, flux, direction, case, sense):
    """SLAU (Shima & Kitamura 2011) with |M+-| >= 1 on both sides (M relative to c_bar): Beta = 1 / 0
    so p_tilde = p of the upwind side; M_hat = 1 so chi = 0 (and theta = 0 for SD-SLAU since
    |M_bar| >= 1); g = 0; the mass flux 0.5 (m_L + m_R - |Vn_bar| d(rho)) with
    Vn_bar = (rho_L |V+| + rho_R |V-|) / (rho_L + rho_R) collapses to rho_up * (m_L + m_R) / (rho_L + rho_R),
    carried with the upwind velocities and total enthalpy."""
    a, b = SUPERSONIC[case]
    L = along(direction, sense * a[1], a[2], a[0], a[3])
    R = along(direction, sense * b[1], b[2], b[0], b[3])
    F = edge_flux(oracle, flux, L, R, direction)
    up = L if sense > 0 else R
    nx, ny = normal(direction)
    vL, vR = L[1] * nx + L[2] * ny, R[1] * nx + R[2] * ny
    mdot = up[0] * (L[0] * vL + R[0] * vR) / (L[0] + R[0])
    H = G / (G - 1.0) * up[3] / up[0] + 0.5 * (up[1] ** 2 + up[2] ** 2)
    ref = np.array([mdot, mdot * up[1] + up[3] * nx, mdot * up[2] + up[3] * ny, mdot * H])
    assert close(F, ref, 1e-14), (F, ref)
    # and not the plain upwind flux whenever the normal mass fluxes differ
    assert not close(F, phys(up, direction), 1e-6)


# ---------------------------------------------------------------------------------------------
# HLLC (Toro 2009 ch. 10, PVRS wave speeds)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("un", [0.0, 0.37, -0.52])
def test_hllc_resolves_isolated_contact_and_shear_exactly(oracle, direction, un):
    """A contact / shear wave (equal normal velocity and pressure, jumps in density and tangential
    velocity) is an exact solution moving at un: HLLC returns the upwind flux (the left one for
    un >= 0), where HLL would smear it by O(jump)."""
    L = along(direction, un, 0.4, 1.3, 0.8)
    R = along(direction, un, -0.7, 0.25, 0.8)
    F = edge_flux(oracle, "HLLC", L, R, direction)
    ref = phys(L if un >= 0.0 else R, direction)
    assert close(F, ref, 1e-14), (F, ref)


def toro_hllc_alternative(L, R, direction):
    """Toro (2009) eq. 10.41-10.42: F*_K = (S*(S_K U_K - F_K) + S_K p*_K D*) / (S_K - S*),
    p*_K = p_K + rho_K (S_K - q_K)(S* - q_K), D* = (0, n_x, n_y, S*), with the PVRS pressure estimate
    and pressure-based wave speeds (eq. 10.59-10.61, 10.67-10.70)."""
    nx, ny = normal(direction)

    def parts(S):
        rho, u, v, p = S
        c = math.sqrt(G * p / rho)
        q = u * nx + v * ny
        E = p / (G - 1.0) + 0.5 * rho * (u * u + v * v)
        U = np.array([rho, rho * u, rho * v, E])
        return rho, u, v, p, c, q, U, phys(S, direction)

    rL, _, _, pL, cL, qL, UL, FL = parts(L)
    rR, _, _, pR, cR, qR, UR, FR = parts(R)
    ppv = max(0.0, 0.5 * (pL + pR) - 0.5 * (qR - qL) * 0.5 * (rL + rR) * 0.5 * (cL + cR))

    def qk(pk):
        return 1.0 if ppv <= pk else math.sqrt(1.0 + (G + 1.0) / (2.0 * G) * (ppv / pk - 1.0))

    SL, SR = qL - cL * qk(pL), qR + cR * qk(pR)
    Ss = (pR - pL + rL * qL * (SL - qL) - rR * qR * (SR - qR)) / (rL * (SL - qL) - rR * (SR - qR))
    if SL >= 0.0:
        return FL
    if SR <= 0.0:
        return FR
    D = np.array([0.0, nx, ny, Ss])
    if Ss >= 0.0:
        pst = pL + rL * (SL - qL) * (Ss - qL)
        return (Ss * (SL * UL - FL) + SL * pst * D) / (SL - Ss)
    pst = pR + rR * (SR - qR) * (Ss - qR)
    return (Ss * (SR * UR - FR) + SR * pst * D) / (SR - Ss)


@pytest.mark.parametrize("direction", [0, 1])
def test_hllc_matches_toro_alternative_star_flux(oracle, direction):
    rng = np.random.default_rng(41)
    n = 0
    for _ in range(200):
        L = (rng.uniform(0.1, 2.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(0.1, 1.5))
        R = (rng.uniform(0.1, 2.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(0.1, 1.5))
        F = edge_flux(oracle, "HLLC", L, R, direction)
        ref = toro_hllc_alternative(L, R, direction)
        assert close(F, ref, 1e-12), (L, R, F, ref)
        n += 1
    assert n == 200


# ---------------------------------------------------------------------------------------------
# pressure-split polynomials, extracted from the momentum flux
# ---------------------------------------------------------------------------------------------
def P_plus(M, alpha):
    return 0.25 * (M + 1.0) ** 2 * (2.0 - M) + alpha * M * (M * M - 1.0) ** 2


@pytest.mark.parametrize("flux,alpha", [("AUSMPW+", 1.0 / 8.0), ("M-AUSMPW+", 0.0), ("AUSM+", 3.0 / 16.0)])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("M", [-0.8, -0.35, 0.05, 0.3, 0.75])
def test_pressure_split_polynomial_and_its_alpha(oracle, flux, alpha, direction, M):
    """Equal normal velocity U and equal temperature p/rho on both sides make the interface sound
    speed a closed form: c_s = sqrt(2 (g-1)/(g+1) H_n) for AUSMPW+ / M-AUSMPW+ (H_n the total
    enthalpy less the tangential kinetic energy) and a* = sqrt(2 (g-1)/(g+1) H) for AUSM+; subsonic
    |U| < c gives M = U / c.  The momentum flux less (mass flux) * U is then P+(M) p_L + P-(M) p_R with
    P+ + P- = 1, so P+(M) = (pressure flux - p_R) / (p_L - p_R), which must be Liou's polynomial with
    the scheme's alpha -- and must not be it with another alpha."""
    T, vt = 0.9, 0.2
    H_n = G / (G - 1.0) * T
    gf = 2.0 * (G - 1.0) / (G + 1.0)
    # AUSM+'s a* uses the whole total enthalpy, AUSMPW+'s c_s the part without the tangential velocity;
    # U = M * c(U) solved for U
    H0 = H_n + (0.5 * vt * vt if flux == "AUSM+" else 0.0)
    U = M * math.sqrt(gf * H0 / (1.0 - 0.5 * gf * M * M))
    c = math.sqrt(gf * (H_n + 0.5 * U * U + (0.5 * vt * vt if flux == "AUSM+" else 0.0)))
    assert abs(U / c - M) < 1e-14
    pL, pR = 1.0, 0.55
    L = along(direction, U, vt, pL / T, pL)
    R = along(direction, U, vt, pR / T, pR)
    F = edge_flux(oracle, flux, L, R, direction)
    k = 1 if direction == 0 else 2
    pressure_flux = F[k] - F[0] * U
    got = (pressure_flux - pR) / (pL - pR)
    assert abs(got - P_plus(M, alpha)) <= 1e-12, (got, P_plus(M, alpha))
    for other in (0.0, 1.0 / 8.0, 3.0 / 16.0):
        if other != alpha and M * (M * M - 1.0) ** 2 != 0.0:
            assert abs(got - P_plus(M, other)) > 1e-6
    # the tangential momentum carries no pressure
    assert abs(F[3 - k] - F[0] * vt) <= 1e-13 * max(abs(F[0] * vt), 1.0)


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("ML,MR", [(0.6, -0.3), (0.2, 0.7), (-0.5, -0.1)])
def test_ausm_plus_mach_split_beta(oracle, direction, ML, MR):
    """AUSM+ interface Mach number M+(M_L) + M-(M_R) with Liou's beta = 1/8: temperatures chosen so that
    both sides have the same total enthalpy make a_half = a* a closed form (subsonic sides), and the
    mass flux is M_half * a* * rho_upwind."""
    gf = 2.0 * (G - 1.0) / (G + 1.0)
    H = 4.0
    a = math.sqrt(gf * H)
    UL, UR = ML * a, MR * a
    TL = (H - 0.5 * UL * UL) * (G - 1.0) / G
    TR = (H - 0.5 * UR * UR) * (G - 1.0) / G
    pL, pR = 1.0, 0.8
    L = along(direction, UL, 0.0, pL / TL, pL)
    R = along(direction, UR, 0.0, pR / TR, pR)
    F = edge_flux(oracle, "AUSM+", L, R, direction)

    def Mp(M, b):
        return 0.25 * (M + 1.0) ** 2 + b * (M * M - 1.0) ** 2

    def Mm(M, b):
        return -0.25 * (M - 1.0) ** 2 - b * (M * M - 1.0) ** 2

    Mh = Mp(ML, 1.0 / 8.0) + Mm(MR, 1.0 / 8.0)
    rho_up = L[0] if Mh > 0 else R[0]
    assert abs(F[0] - Mh * a * rho_up) <= 1e-13 * abs(F[0]), (F[0], Mh * a * rho_up)
    wrong = Mp(ML, 0.0) + Mm(MR, 0.0)
    assert abs(F[0] - wrong * a * rho_up) > 1e-6


# ---------------------------------------------------------------------------------------------
# the reference's Roe variant
# ---------------------------------------------------------------------------------------------
def roe_averages(L, R):
    rl, ul, vl, pl = L
    rr, ur, vr, pr = R
    hl = G / (G - 1.0) * pl / rl + 0.5 * (ul * ul + vl * vl)
    hr = G / (G - 1.0) * pr / rr + 0.5 * (ur * ur + vr * vr)
    wl, wr = math.sqrt(rl), math.sqrt(rr)
    u = (wl * ul + wr * ur) / (wl + wr)
    v = (wl * vl + wr * vr) / (wl + wr)
    h = (wl * hl + wr * hr) / (wl + wr)
    c = math.sqrt((G - 1.0) * (h - 0.5 * (u * u + v * v)))
    return math.sqrt(rl * rr), u, v, h, c, hl, hr


@pytest.mark.parametrize("direction", [0, 1])
def test_roe_variant_offset_in_supersonic_flow(oracle, direction):
    """All four waves positive with the reference's wave-speed clips inactive (u~ - c~ <= q_L - c_L,
    u~ + c~ >= q_R + c_R) and no pressure sensor: the dissipation is exactly |A~| dU = A~ dU =
    F(U_R) - F(U_L) (Roe property), so 0.5 (F_L(u~) + F_R(u~) - A~ dU) = F(U_L) + 0.5 [rho_L (u~ - q_L)
    (1, u_L, v_L, h_L) + rho_R (u~ - q_R)(1, u_R, v_R, h_R)]: zero for textbook Roe, which evaluates F_L
    and F_R with q_L and q_R."""
    L = along(direction, 3.2, 0.3, 1.0, 1.0)
    R = along(direction, 3.0, -0.2, 1.1, 1.05)
    F = edge_flux(oracle, "Roe", L, R, direction)
    nx, ny = normal(direction)
    _, ua, va, _, ca, hl, hr = roe_averages(L, R)
    uv = ua * nx + va * ny
    qL, qR = L[1] * nx + L[2] * ny, R[1] * nx + R[2] * ny
    cL, cR = math.sqrt(G * L[3] / L[0]), math.sqrt(G * R[3] / R[0])
    assert uv - ca > 0.0 and uv - ca <= qL - cL and uv + ca >= qR + cR  # the case the closed form needs
    off = 0.5 * (L[0] * (uv - qL) * np.array([1.0, L[1], L[2], hl]) +
                 R[0] * (uv - qR) * np.array([1.0, R[1], R[2], hr]))
    ref = phys(L, direction) + off
    assert close(F, ref, 1e-13), (F, ref)
    assert np.max(np.abs(off)) > 1e-3  # the variant is visible here


@pytest.mark.parametrize("direction", [0, 1])
def test_roe_pressure_sensor_correction_term(oracle, direction):
    """roe_solver.fypp:495-511: once a pressure ratio in the sensor stencil reaches 2 (f_h = 1), the
    momentum components gain -0.5 rho~ c~ (-dq_t n_y + psi dq_n n_x, dq_t n_x + psi dq_n n_y), with
    psi = |c_R - c_L| / max(c_R, c_L); mass and energy are unchanged."""
    L = along(direction, 0.3, 0.5, 1.0, 1.0)
    R = along(direction, -0.2, -0.4, 0.7, 0.8)
    F0 = edge_flux(oracle, "Roe", L, R, direction)
    p = np.ones((NI + 4, NJ + 4), order="F")
    p[I0 + 1, J0 + 1] = 3.0  # cell (i, j) itself
    F1 = edge_flux(oracle, "Roe", L, R, direction, pcell=p)
    nx, ny = normal(direction)
    rav, ua, va, _, ca, _, _ = roe_averages(L, R)
    dqn = (R[1] - L[1]) * nx + (R[2] - L[2]) * ny
    dqt = -((R[1] - L[1]) * ny) + (R[2] - L[2]) * nx
    cL, cR = math.sqrt(G * L[3] / L[0]), math.sqrt(G * R[3] / R[0])
    psi = abs(cR - cL) / max(cR, cL)
    d = np.array([0.0, -0.5 * rav * ca * (-(dqt * ny) + psi * dqn * nx),
                  -0.5 * rav * ca * (dqt * nx + psi * dqn * ny), 0.0])
    assert np.max(np.abs((F1 - F0) - d)) <= 1e-13 * max(np.max(np.abs(F0)), 1.0), (F1 - F0, d)
    assert np.max(np.abs(d)) > 1e-3


# ---------------------------------------------------------------------------------------------
# the cells each shock sensor reads
# ---------------------------------------------------------------------------------------------
# offsets (di, dj) from cell (i, j) on the left / bottom of the edge, read from the reference
SENSOR_STENCILS = {
    # schemes that read no cell pressure at all
    "SLAU": (set(), set()),
    "SLAU2": (set(), set()),
    "AUSM+": (set(), set()),
    "HLLC": (set(), set()),
    # slau_solver.fypp:413-429 (the pressures entering delta_p1..delta_p6)
    "SD-SLAU": ({(0, 1), (0, 0), (-1, 0), (0, -1), (1, -1), (1, 0), (2, 0), (1, 1)},
                {(-1, 0), (0, 0), (0, -1), (1, 0), (1, 1), (0, 1), (0, 2), (-1, 1)}),
    # ausmpw_plus_solver.fypp:432-441 (get_f's p1_L, p1_R, p2_L, p2_R)
    "AUSMPW+": ({(0, -1), (1, -1), (0, 1), (1, 1)}, {(1, 0), (1, 1), (-1, 0), (-1, 1)}),
    # m_ausmpw_plus_solver.fypp:672-681
    "M-AUSMPW+": ({(0, -1), (1, -1), (0, 1), (1, 1)}, {(1, 0), (1, 1), (-1, 0), (-1, 1)}),
    # roe_solver.fypp:495-507 (the cells of f_p(1..5))
    "Roe": ({(0, 0), (1, 0), (0, 1), (0, -1), (1, 1), (1, -1)},
            {(0, 0), (0, 1), (-1, 0), (1, 0), (-1, 1), (1, 1)}),
}
SENSOR_STENCILS["SD-SLAU2"] = SENSOR_STENCILS["SD-SLAU"]


@pytest.mark.parametrize("flux", sorted(SENSOR_STENCILS))
@pytest.mark.parametrize("direction", [0, 1])
def test_shock_sensor_stencil(oracle, flux, direction):
    """Raising (x3) or lowering (x0.3) the pressure of one cell changes the edge's flux exactly when that
    cell belongs to the scheme's sensor stencil."""
    L = along(direction, 0.15, 0.1, 1.0, 1.0)
    R = along(direction, 0.1, -0.05, 0.9, 1.12)
    base = edge_flux(oracle, flux, L, R, direction)
    seen = set()
    for di in range(-3, 4):
        for dj in range(-3, 4):
            for fac in (3.0, 0.3):
                p = np.ones((NI + 4, NJ + 4), order="F")
                p[I0 + 1 + di, J0 + 1 + dj] = fac  # halo'd index of cell (i + di, j + dj)
                F = edge_flux(oracle, flux, L, R, direction, pcell=p)
                if not np.array_equal(F, base):
                    seen.add((di, dj))
    assert seen == SENSOR_STENCILS[flux][direction], sorted(seen)


# ---------------------------------------------------------------------------------------------
# error codes
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", sorted(ERROR_CASES))
def test_error_codes_follow_the_reference_check_order(oracle, name):
    """NaN in any stage's rho, u, v or p stops first (assign_fluid, NANS_FOUND = 3); after the step
    negative density (NEG_DENSITY = 1) is checked before negative pressure (NEG_PRESSURE = 2)."""
    code, f = error_case_oracle(oracle, name)
    expect = ERROR_CASES[name]["code"]
    assert code == expect
    rho, p = f.interior("rho"), f.interior("p")
    nan = bool(np.isnan(rho).any() or np.isnan(p).any())
    if expect == 3:
        assert nan
    else:
        assert not nan
        assert (expect == 1) == bool((rho < 0).any())
        assert (expect == 2) == (not (rho < 0).any() and bool((p < 0).any()))
    if ERROR_CASES[name].get("both"):
        assert (rho < 0).any() and (p < 0).any()
