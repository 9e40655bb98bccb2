"""Exact solution of the 1D Riemann problem for an ideal gas.

TEST INFRASTRUCTURE ONLY: a closed-form checker for the fluid sweeps.

Written from the published algorithm of E. F. Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4 (pressure function
f_K(p), root of f_L + f_R + du = 0, sampling of the self-similar solution).
The reference computes the same thing in MATLAB for its Sod convergence test
(experiment/RiemannProblem/analysis/SodShockSolution.m:32-277, read for its
conventions only: cell averages, interface position).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.optimize import brentq


def _fk(p, rho, pk, ck, g):
    """Toro eq. 4.6/4.7: pressure function and its derivative for one side."""
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        s = math.sqrt(a / (p + b))
        return (p - pk) * s
    return 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)


def star_state(left, right, gamma):
    """Returns (p_star, u_star) for left/right primitive states (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl = math.sqrt(gamma * pl / rl)
    cr = math.sqrt(gamma * pr / rr)
    du = ur - ul
    # pressure positivity condition (Toro eq. 4.40)
    if 2.0 * (cl + cr) / (gamma - 1.0) <= du:
        raise ValueError("vacuum is generated; not supported")

    def fn(p):
        return _fk(p, rl, pl, cl, gamma) + _fk(p, rr, pr, cr, gamma) + du

    lo = 1e-14 * min(pl, pr)
    hi = max(pl, pr)
    while fn(hi) < 0:
        hi *= 2.0
    ps = brentq(fn, lo, hi, xtol=1e-15 * hi, rtol=1e-15, maxiter=500)
    us = 0.5 * (ul + ur) + 0.5 * (_fk(ps, rr, pr, cr, gamma) - _fk(ps, rl, pl, cl, gamma))
    return ps, us


def sample(xi, left, right, gamma, ps=None, us=None):
    """Primitive state (rho, u, p) at similarity coordinate xi = x/t (Toro 4.5)."""
    rl, ul, pl = left
    rr, ur, pr = right
    g = gamma
    if ps is None:
        ps, us = star_state(left, right, gamma)
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    if xi <= us:  # left of contact
        if ps > pl:  # left shock
            sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
            if xi <= sl:
                return rl, ul, pl
            r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
            return r, us, ps
        # left rarefaction
        shl = ul - cl
        if xi <= shl:
            return rl, ul, pl
        cml = cl * (ps / pl) ** ((g - 1) / (2 * g))
        stl = us - cml
        if xi > stl:
            return rl * (ps / pl) ** (1 / g), us, ps
        u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + xi)
        c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - xi))
        r = rl * (c / cl) ** (2 / (g - 1))
        p = pl * (c / cl) ** (2 * g / (g - 1))
        return r, u, p
    # right of contact
    if ps > pr:  # right shock
        sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
        if xi >= sr:
            return rr, ur, pr
        r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
        return r, us, ps
    shr = ur + cr
    if xi >= shr:
        return rr, ur, pr
    cmr = cr * (ps / pr) ** ((g - 1) / (2 * g))
    str_ = us + cmr
    if xi <= str_:
        return rr * (ps / pr) ** (1 / g), us, ps
    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + xi)
    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - xi))
    r = rr * (c / cr) ** (2 / (g - 1))
    p = pr * (c / cr) ** (2 * g / (g - 1))
    return r, u, p


def cell_averages(xc, dx, x0, t, left, right, gamma, nsub=32):
    """Cell-averaged conserved (rho, mom, E) over cells centred at xc (width dx),
    with the initial discontinuity at x0, at time t (midpoint sub-sampling)."""
    ps, us = star_state(left, right, gamma)
    xc = np.asarray(xc, dtype=np.float64)
    off = (np.arange(nsub) + 0.5) / nsub - 0.5
    out = np.zeros((3, xc.size))
    for k, x in enumerate(xc):
        acc = np.zeros(3)
        for o in off:
            xi = (x + o * dx - x0) / t
            r, u, p = sample(xi, left, right, gamma, ps, us)
            acc += (r, r * u, p / (gamma - 1) + 0.5 * r * u * u)
        out[:, k] = acc / nsub
    return out
