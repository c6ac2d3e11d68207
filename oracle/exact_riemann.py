"""Exact solution of the 1-D Riemann problem for an ideal gas (TEST INFRASTRUCTURE).

Textbook algorithm (E. F. Toro, "Riemann Solvers and Numerical Methods for
Fluid Dynamics", ch. 4): Newton iteration on the pressure function
f(p) = f_L(p) + f_R(p) + u_R - u_L for the star pressure, then sampling of
the self-similar solution at x/t.  Used to pin the Sod shock tube
(ic/shocktube.f90) independently of the reference code.
"""
from __future__ import annotations

import math

import numpy as np


def _fk(p, rho, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(a / (p + b))
        f = (p - pk) * q
        df = q * (1.0 - 0.5 * (p - pk) / (b + p))
    else:  # rarefaction
        pr = p / pk
        f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    """(p*, u*) by Newton iteration from the two-rarefaction guess."""
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    z = (g - 1.0) / (2.0 * g)
    p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
    for _ in range(100):
        fl, dfl = _fk(p, rl, pl, cl, g)
        fr, dfr = _fk(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(p - dp, 1e-14)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(rl, ul, pl, rr, ur, pr, g, xi):
    """rho, u, p at similarity coordinates xi = (x - x_diss)/t (array)."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    xi = np.asarray(xi, dtype=np.float64)
    rho, u, p = np.empty_like(xi), np.empty_like(xi), np.empty_like(xi)
    for i, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    rho[i], u[i], p[i] = rl, ul, pl
                else:
                    r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                    rho[i], u[i], p[i] = r, us, ps
            else:  # left rarefaction
                shl = ul - cl
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - csl
                if s <= shl:
                    rho[i], u[i], p[i] = rl, ul, pl
                elif s >= stl:
                    rho[i], u[i], p[i] = rl * (ps / pl) ** (1 / g), us, ps
                else:
                    uu = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    rho[i] = rl * (c / cl) ** (2 / (g - 1))
                    u[i] = uu
                    p[i] = pl * (c / cl) ** (2 * g / (g - 1))
        else:  # right of contact
            if ps > pr:  # right shock
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    rho[i], u[i], p[i] = rr, ur, pr
                else:
                    r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                    rho[i], u[i], p[i] = r, us, ps
            else:  # right rarefaction
                shr = ur + cr
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + csr
                if s >= shr:
                    rho[i], u[i], p[i] = rr, ur, pr
                elif s <= str_:
                    rho[i], u[i], p[i] = rr * (ps / pr) ** (1 / g), us, ps
                else:
                    uu = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    rho[i] = rr * (c / cr) ** (2 / (g - 1))
                    u[i] = uu
                    p[i] = pr * (c / cr) ** (2 * g / (g - 1))
    return rho, u, p
