"""Exact solution of the 1D Riemann problem for an ideal gas (Toro, "Riemann
Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4: pressure
function f_K of eq. 4.6-4.7, Newton iteration 4.44, sampling of sec. 4.5).
Used only as an accuracy check (not parity) for the Sod shock tube."""
import math

import numpy as np


def _fk(p, rk, pk, ck, g):
    if p > pk:
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        f = (p - pk) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1.0 - (p - pk) / (2.0 * (B + p)))
    else:
        f = 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rk * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g=1.4):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dfl = _fk(p, rl, pl, cl, g)
        fr, dfr = _fk(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-12)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, rl, ul, pl, rr, ur, pr, g=1.4):
    """Density, velocity, pressure at positions x and time t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((3, len(x)))
    for k, xx in enumerate(x):
        s = (xx - x0) / t
        if s <= us:
            if ps > pl:
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    r, u, p = rl, ul, pl
                else:
                    r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                    u, p = us, ps
            else:
                shl = ul - cl
                cml = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cml
                if s <= shl:
                    r, u, p = rl, ul, pl
                elif s > stl:
                    r, u, p = rl * (ps / pl) ** (1 / g), us, ps
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    p = pl * (c / cl) ** (2 * g / (g - 1))
        else:
            if ps > pr:
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    r, u, p = rr, ur, pr
                else:
                    r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                    u, p = us, ps
            else:
                shr = ur + cr
                cmr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + cmr
                if s >= shr:
                    r, u, p = rr, ur, pr
                elif s <= str_:
                    r, u, p = rr * (ps / pr) ** (1 / g), us, ps
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    p = pr * (c / cr) ** (2 * g / (g - 1))
        out[:, k] = (r, u, p)
    return out
