"""Independent textbook formulations used to pin the oracle (no reference code).

* HLLC flux in conserved-variable form, F*_K = F_K + S_K (U*_K - U_K)
  (Toro, Riemann Solvers, 3rd ed., eqs. 10.37-10.39), with Einfeldt's HLLE
  wave-speed estimates from Roe averages of velocity and total enthalpy
  (Einfeldt 1988, SIAM J. Numer. Anal. 25, 294).
* A first-order (piecewise-constant) Heun / SSP-RK2 line update built on it.
"""
import numpy as np


def prim(U, g):
    rho = U[0]
    v = U[2:5] / rho
    p = (g - 1) * (U[1] - 0.5 * rho * (v ** 2).sum(axis=0))
    return rho, v, p


def phys_flux(U, g):
    """Euler flux along the first velocity component. U rows (rho, E, mN, mT1, mT2)."""
    rho, v, p = prim(U, g)
    u = v[0]
    F = np.empty_like(U)
    F[0] = rho * u
    F[1] = (U[1] + p) * u
    F[2] = U[2] * u + p
    F[3] = U[3] * u
    F[4] = U[4] * u
    return F


def hllc_textbook(UL, UR, g):
    rl, vl, pl = prim(UL, g)
    rr, vr, pr = prim(UR, g)
    ul, ur = vl[0], vr[0]
    HL = (UL[1] + pl) / rl
    HR = (UR[1] + pr) / rr
    sl_, sr_ = np.sqrt(rl), np.sqrt(rr)
    vt = (sl_ * vl + sr_ * vr) / (sl_ + sr_)
    Ht = (sl_ * HL + sr_ * HR) / (sl_ + sr_)
    ct = np.sqrt((g - 1) * (Ht - 0.5 * (vt ** 2).sum(axis=0)))
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    SL = np.minimum(ul - cl, vt[0] - ct)
    SR = np.maximum(ur + cr, vt[0] + ct)
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))

    def ustar(U, r, v, p, S):
        u = v[0]
        f = r * (S - u) / (S - Ss)
        out = np.empty_like(U)
        out[0] = f
        out[1] = f * (U[1] / r + (Ss - u) * (Ss + p / (r * (S - u))))
        out[2] = f * Ss
        out[3] = f * v[1]
        out[4] = f * v[2]
        return out

    FL = phys_flux(UL, g)
    FR = phys_flux(UR, g)
    FsL = FL + SL * (ustar(UL, rl, vl, pl, SL) - UL)
    FsR = FR + SR * (ustar(UR, rr, vr, pr, SR) - UR)
    return np.where(SL >= 0, FL, np.where(Ss >= 0, FsL, np.where(SR > 0, FsR, FR)))


def first_order_heun_line(U, lam, g):
    """SSP-RK2 with piecewise-constant states on a periodic line. U: (5, n)."""
    def dF(V):
        F = hllc_textbook(V, np.roll(V, -1, axis=1), g)   # flux at i+1/2
        return F - np.roll(F, 1, axis=1)

    U1 = U - lam * dF(U)
    return 0.5 * (U1 + U - lam * dF(U1))


def einfeldt_speeds(UL, UR, g):
    """Einfeldt's HLLE bounds from Roe averages of velocity and enthalpy."""
    rl, vl, pl = prim(UL, g)
    rr, vr, pr = prim(UR, g)
    HL = (UL[1] + pl) / rl
    HR = (UR[1] + pr) / rr
    sl_, sr_ = np.sqrt(rl), np.sqrt(rr)
    vt = (sl_ * vl + sr_ * vr) / (sl_ + sr_)
    Ht = (sl_ * HL + sr_ * HR) / (sl_ + sr_)
    ct = np.sqrt((g - 1) * (Ht - 0.5 * (vt ** 2).sum(axis=0)))
    SL = np.minimum(vl[0] - np.sqrt(g * pl / rl), vt[0] - ct)
    SR = np.maximum(vr[0] + np.sqrt(g * pr / rr), vt[0] + ct)
    return SL, SR


def hlle_textbook(UL, UR, g):
    """HLL flux (Toro eq. 10.21) with Einfeldt speeds."""
    SL, SR = einfeldt_speeds(UL, UR, g)
    FL = phys_flux(UL, g)
    FR = phys_flux(UR, g)
    Fm = (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)
    return np.where(SL >= 0, FL, np.where(SR <= 0, FR, Fm))


def first_order_heun_line_with(flux, U, lam, g):
    """SSP-RK2 with piecewise-constant states and the given two-state flux."""
    def dF(V):
        F = flux(V, np.roll(V, -1, axis=1), g)
        return F - np.roll(F, 1, axis=1)

    U1 = U - lam * dF(U)
    return 0.5 * (U1 + U - lam * dF(U1))


def rusanov_step(U, lam, g, c):
    """One forward-Euler step with the local Lax-Friedrichs (Rusanov) flux
    G_{i+1/2} = (F_i + F_{i+1})/2 - c (U_{i+1} - U_i)/2 and a fixed speed c."""
    F = phys_flux(U, g)
    Up, Fp = np.roll(U, -1, axis=1), np.roll(F, -1, axis=1)
    G = 0.5 * (F + Fp) - 0.5 * c * (Up - U)
    return U - lam * (G - np.roll(G, 1, axis=1))
