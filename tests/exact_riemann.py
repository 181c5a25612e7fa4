"""Exact solution of the 1D Riemann problem for an ideal gas (Toro, Riemann Solvers and Numerical
Methods for Fluid Dynamics, ch. 4): Newton iteration on the pressure function for p*, then
sampling of the self-similar solution.  Test helper (physics oracle for Sod)."""
import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1 - 0.5 * (p - pk) / (B + p))
    pr = p / pk  # rarefaction
    return 2 * ck / (g - 1) * (pr ** ((g - 1) / (2 * g)) - 1), 1.0 / (rho * ck) * pr ** (-(g + 1) / (2 * g))


def star_state(L, R, g=1.4):
    (rl, ul, pl), (rr, ur, pr) = L, R
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        p_new = max(1e-12, p - dp)
        if abs(p_new - p) < 1e-15 * p_new:
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, L, R, g=1.4):
    """Primitive (rho, u, p) at positions x, time t."""
    (rl, ul, pl), (rr, ur, pr) = L, R
    ps, us = star_state(L, R, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((3, len(x)))
    for k, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                st = (rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)), us, ps)
                out[:, k] = (rl, ul, pl) if s <= sl else st
            else:
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                shl, stl = ul - cl, us - csl
                if s <= shl:
                    out[:, k] = (rl, ul, pl)
                elif s >= stl:
                    out[:, k] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    rho = rl * (c / cl) ** (2 / (g - 1))
                    out[:, k] = (rho, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                st = (rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)), us, ps)
                out[:, k] = (rr, ur, pr) if s >= sr else st
            else:
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                shr, str_ = ur + cr, us + csr
                if s >= shr:
                    out[:, k] = (rr, ur, pr)
                elif s <= str_:
                    out[:, k] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    rho = rr * (c / cr) ** (2 / (g - 1))
                    out[:, k] = (rho, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
