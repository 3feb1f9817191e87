"""Exact solution of the 1-D Riemann problem for an ideal gas.

Textbook algorithm (E. F. Toro, Riemann Solvers and Numerical Methods for
Fluid Dynamics, 3rd ed., ch. 4): Newton iteration on the pressure function
f(p) = f_L(p) + f_R(p) + (u_R - u_L) for the star pressure, then sampling of
the self-similar solution at x/t.  Independent of the solver under test; it
pins the HLLC/MUSCL/RK path in the single-material limit.
"""
import math

import numpy as np


def _fk(p, rho, pk, g):
    a = math.sqrt(g * pk / rho)
    if p > pk:   # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    r = (p / pk) ** ((g - 1.0) / (2.0 * g))
    return 2.0 * a / (g - 1.0) * (r - 1.0), 1.0 / (rho * a) * (p / pk) ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g):
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _fk(p, rl, pl, g)
        fr, dr = _fk(p, rr, pr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(1e-12, p - dp)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, g)
    fr, _ = _fk(p, rr, pr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, rl, ul, pl, rr, ur, pr, g):
    """(rho, u, p) at the points x at time t for the discontinuity at x0."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    out = np.zeros((len(x), 3))
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:   # left of contact
            if ps > pl:
                sl = ul - al * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    out[i] = (rl, ul, pl)
                else:
                    r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                    out[i] = (r, us, ps)
            else:
                shl = ul - al
                if s <= shl:
                    out[i] = (rl, ul, pl)
                else:
                    ast = al * (ps / pl) ** ((g - 1) / (2 * g))
                    stl = us - ast
                    if s > stl:
                        out[i] = (rl * (ps / pl) ** (1 / g), us, ps)
                    else:
                        c = 2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)
                        out[i] = (rl * c ** (2 / (g - 1)), 2 / (g + 1) * (al + (g - 1) / 2 * ul + s),
                                  pl * c ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + ar * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    out[i] = (rr, ur, pr)
                else:
                    r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                    out[i] = (r, us, ps)
            else:
                shr = ur + ar
                if s >= shr:
                    out[i] = (rr, ur, pr)
                else:
                    ast = ar * (ps / pr) ** ((g - 1) / (2 * g))
                    str_ = us + ast
                    if s <= str_:
                        out[i] = (rr * (ps / pr) ** (1 / g), us, ps)
                    else:
                        c = 2 / (g + 1) - (g - 1) / ((g + 1) * ar) * (ur - s)
                        out[i] = (rr * c ** (2 / (g - 1)), 2 / (g + 1) * (-ar + (g - 1) / 2 * ur + s),
                                  pr * c ** (2 * g / (g - 1)))
    return out
