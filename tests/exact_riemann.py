"""Exact solution of the 1D Euler Riemann problem for an ideal gas.

Textbook exact solver (Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", ch. 4): Newton iteration on the star pressure with the shock and
rarefaction branches of the pressure function, then sampling of the
self-similar solution at x/t.  Used to pin the whole path against the
reference testsuite's Sod error norms (it is not the reference's code).
"""
from __future__ import annotations

import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(a / (p + b))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (b + p))
    pr = p / pk  # rarefaction
    f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(1e-14, p - dp)
        if abs(pn - p) <= 1e-15 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(xi, rl, ul, pl, rr, ur, pr, g):
    """Density, velocity, pressure at similarity coordinates xi = (x-x0)/t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    out = np.empty((3, len(xi)))
    for n, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pl:
                rs = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                out[:, n] = (rl, ul, pl) if s <= sl else (rs, us, ps)
            else:
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                head, tail = ul - cl, us - cs
                if s <= head:
                    out[:, n] = (rl, ul, pl)
                elif s >= tail:
                    out[:, n] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[:, n] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                rs = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                out[:, n] = (rr, ur, pr) if s >= sr else (rs, us, ps)
            else:
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                head, tail = ur + cr, us + cs
                if s >= head:
                    out[:, n] = (rr, ur, pr)
                elif s <= tail:
                    out[:, n] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[:, n] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
