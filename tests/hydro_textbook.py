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


# ---------------------------------------------------------------- limiters, r-form
# Half-slopes s(A, B) = B psi(A/B) / 2 from the textbook limiter functions
# psi(r) of the slope ratio r = A/B (Sweby 1984; Waterson & Deconinck 2007 for
# Ospre; van Albada et al. 1982), with psi = 0 for r <= 0.
def psi_ospre(r):
    return np.where(r > 0, 1.5 * (r * r + r) / (r * r + r + 1), 0.0)


def psi_vanleer(r):
    return np.where(r > 0, 2 * r / (1 + r), 0.0)


def psi_vanalbada(r):
    return np.where(r >= 0, (r * r + r) / (r * r + 1), 0.0)


def psi_mc(r):
    return np.where(r > 0, np.minimum(np.minimum(0.5 * (1 + r), 2 * r), 2.0), 0.0)


def half_slope(psi, A, B):
    """s = B psi(A/B) / 2 elementwise; 0 where B == 0 (both forms vanish there)."""
    A = np.asarray(A, dtype=float)
    B = np.asarray(B, dtype=float)
    safe = np.where(B == 0, 1.0, B)
    return np.where(B == 0, 0.0, 0.5 * B * psi(A / safe))


# ---------------------------------------------------------------- MUSCL + SSP-RK2
def to_cons(rho, e, v):
    """Conserved (rho, E, m) from density, internal energy density e and velocity (3, n)."""
    U = np.empty((5,) + np.shape(rho))
    U[0] = rho
    U[1] = e + 0.5 * rho * (v ** 2).sum(axis=0)
    U[2:5] = rho * v
    return U


def muscl_faces(U, g, slope, recon="prim"):
    """Left and right face states (conserved) of every cell of a periodic line.

    recon "prim": limited slopes of density, internal energy density and
    velocity (the variables a MUSCL scheme reconstructs when velocity, not
    momentum, is the reconstruction variable); "cons": of the conserved
    variables themselves; "prim_p_rho": density, velocity and the specific
    internal energy e/rho. slope(A, B) is the half-slope of backward/forward
    differences A, B."""
    if recon == "cons":
        W = U
    else:
        rho, v, p = prim(U, g)
        e = p / (g - 1)
        W = np.vstack([rho[None], (e / rho if recon == "prim_p_rho" else e)[None], v])
    A = W - np.roll(W, 1, axis=1)
    B = np.roll(W, -1, axis=1) - W
    s = slope(A, B)
    WL, WR = W - s, W + s

    def back(X):
        if recon == "cons":
            return X
        eint = X[1] * X[0] if recon == "prim_p_rho" else X[1]
        return to_cons(X[0], eint, X[2:5])

    return back(WL), back(WR)


def muscl_heun_line(flux, U, lam, g, slope, recon="prim"):
    """Second-order MUSCL + SSP-RK2 (Heun) on a periodic line with the
    two-state flux `flux(UL, UR, g)`: interface i+1/2 takes the right face of
    cell i and the left face of cell i+1."""
    def dF(V):
        L, R = muscl_faces(V, g, slope, recon)
        F = flux(R, np.roll(L, -1, axis=1), g)
        return F - np.roll(F, 1, axis=1)

    U1 = U - lam * dF(U)
    return 0.5 * (U1 + U - lam * dF(U1))


# ---------------------------------------------------------------- Jin-Xin relaxation
def jin_xin_line(U, lam, g, psi=psi_ospre, second_order=True, rhomin=0.0, emin=0.0):
    """Relaxed Jin-Xin scheme on a periodic line (Jin & Xin, CPAM 48 (1995)
    235, section 4), written in its interface-flux form with the frozen speed
    c = max_i(|u_i| + c_s,i) over the line:
        G_{i+1/2} = (w+_i + d+_i - w-_{i+1} + d-_{i+1}) / 2,  w+- = c U +- F(U),
    where d+- are MUSCL edge corrections of the characteristic variables
    (d+_i = +slope of w+ towards i+1/2, d-_{i+1} = +slope of w- towards
    i+1/2; zero for the first-order predictor). Time stepping: half-step
    first-order predictor, then the second-order corrector from the
    predicted state (with its own frozen speed), each followed by the density
    and internal-energy floors when rhomin/emin are given."""
    def freeze(V):
        rho, v, p = prim(V, g)
        return np.max(np.abs(v[0]) + np.sqrt(g * p / rho))

    def G(V, c, tvd):
        F = phys_flux(V, g)
        wp, wm = c * V + F, c * V - F
        if tvd:
            def slope(w):
                A = w - np.roll(w, 1, axis=1)
                B = np.roll(w, -1, axis=1) - w
                return half_slope(psi, A, B)
            wp = wp + slope(wp)           # right edge of cell i
            wm = wm - slope(wm)           # left edge of cell i
        return 0.5 * (wp - np.roll(wm, -1, axis=1))

    def floor(V):
        V = V.copy()
        V[0] = np.maximum(V[0], rhomin)
        ke = 0.5 * (V[2] ** 2 + V[3] ** 2 + V[4] ** 2) / V[0]
        low = (V[1] - ke) < V[0] * emin
        V[1] = np.where(low, ke + V[0] * emin, V[1])
        return V

    Gp = G(U, freeze(U), False)
    Uh = floor(U - 0.5 * lam * (Gp - np.roll(Gp, 1, axis=1)))
    if not second_order:
        return Uh
    Gc = G(Uh, freeze(Uh), True)
    return floor(U - lam * (Gc - np.roll(Gc, 1, axis=1)))
