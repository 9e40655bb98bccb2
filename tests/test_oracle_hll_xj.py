"""Pins the oracle's HLL and Xin-Jin sweeps (oracle/imogen_oracle.c) against
textbook formulations, closed forms and invariants. The reference has no
golden vectors for these paths either (run/test_suite/tsSod.m is a run-time
convergence test).

What pins what:
  HLL flux + speeds    textbook HLLE (Toro 10.21, Einfeldt bounds), zero limiter
  HLL floors           closed forms on uniform states (cudaFluidStep.cu:2813-2864)
  XJ predictor         Rusanov half step with the per-line freezing speed
  XJ corrector         2nd-order convergence on a smooth wave (a dropped or
                       mis-signed TVD correction is 1st order or unstable)
  XJ floors            closed form on a uniform state (:2620-2628)
  both                 Sod vs the exact solution, conservation, fixed points,
                       direction maps
  HLL isothermal       the interface flux read off the sweep at a jump
                       (Richardson in lam): upwind physical flux when both
                       bounds share a sign, one HLL speed pair consistent over
                       mass and all momenta when subsonic, and that pair equal
                       to the reference's bounds (C_iso^2 as the sound speed,
                       cs overwritten before the one-sided bounds), with states
                       that separate them from textbook Einfeldt; the
                       energy flux (a passive scalar, the reference's HLL form
                       without 1/(SR-SL), cudaFluidStep.cu:2178-2181) against
                       that form with the fitted speed pair
"""
import numpy as np
import pytest

from oracle import exact_riemann as er
from oracle import pyoracle as po

from . import hydro_textbook as tb

G = 1.4


def cons(rho, vx, vy, vz, P, g=G):
    q = np.zeros((5,) + np.shape(rho))
    q[0] = rho
    q[1] = P / (g - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    q[2] = rho * vx
    q[3] = rho * vy
    q[4] = rho * vz
    return q


def _line(seed, n, vspread=1.5):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.2, 2.0, n)
    v = rng.uniform(-vspread, vspread, (3, n))
    P = rng.uniform(0.1, 3.0, n)
    return cons(rho, v[0], v[1], v[2], P)


# ---------------------------------------------------------------- HLL
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_first_order_hll_matches_textbook_hlle(seed):
    n = 24
    U = _line(seed, n)
    q = U.reshape(5, 1, 1, n).copy()
    po.fluid_update_1d(q, 1, 0.1, G, method=po.HLL, limiter=po.LIM_ZERO, rhomin=1e-12)
    ref = tb.first_order_heun_line_with(tb.hlle_textbook, U, 0.1, G)
    err = np.abs(q.reshape(5, n) - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("vx", [4.0, -4.0])
def test_supersonic_hll_matches_textbook(vx):
    n = 16
    rng = np.random.default_rng(7)
    U = cons(rng.uniform(0.5, 1.5, n), np.full(n, vx), 0.1 * rng.normal(size=n), 0.1 * rng.normal(size=n),
             rng.uniform(0.1, 0.3, n))
    q = U.reshape(5, 1, 1, n).copy()
    po.fluid_update_1d(q, 1, 0.05, G, method=po.HLL, limiter=po.LIM_ZERO, rhomin=1e-12)
    ref = tb.first_order_heun_line_with(tb.hlle_textbook, U, 0.05, G)
    assert np.abs(q.reshape(5, n) - ref).max() / np.abs(ref).max() < 1e-13


def test_hll_density_floor_rescales_the_cell():
    """A uniform state below the density floor: the fluxes cancel, so one sweep
    returns the floored state -- rho = rhomin, E and p scaled by rhomin/rho."""
    rhomin = 1e-3
    q = cons(np.full((3, 4, 8), 2e-4), np.full((3, 4, 8), 0.5), np.full((3, 4, 8), -0.25),
             np.full((3, 4, 8), 0.125), np.full((3, 4, 8), 1e-3))
    f = rhomin / 2e-4
    exp = q * f
    for d in (1, 2, 3):
        r = q.copy()
        po.fluid_update_1d(r, d, 0.1, G, method=po.HLL, rhomin=rhomin)
        assert np.allclose(r, exp, rtol=1e-14, atol=0), d


def test_hll_energy_floor():
    """Internal energy below rho*emin (emin = 1e-5/(gamma-1)) is raised to it."""
    sh = (2, 3, 8)
    rho, vx = 0.7, 1.3
    q = cons(np.full(sh, rho), np.full(sh, vx), np.zeros(sh), np.zeros(sh), np.full(sh, 1e-9))
    r = q.copy()
    po.fluid_update_1d(r, 1, 0.1, G, method=po.HLL, rhomin=1e-8)
    emin = 1e-5 / (G - 1)
    exp_E = 0.5 * rho * vx * vx + rho * emin
    assert np.allclose(r[1], exp_E, rtol=1e-14, atol=0)
    assert np.array_equal(r[[0, 2, 3, 4]], q[[0, 2, 3, 4]])


def test_hll_direction_maps():
    rng = np.random.default_rng(5)
    sh = (7, 9, 11)
    q = cons(rng.uniform(0.5, 2, sh), rng.normal(size=sh), rng.normal(size=sh), rng.normal(size=sh),
             rng.uniform(0.5, 2, sh))

    def sweep(a, d):
        a = a.copy()
        po.fluid_update_1d(a, d, 0.05, G, method=po.HLL)
        return a

    rx = sweep(q, 1)
    qy = np.ascontiguousarray(q.transpose(0, 1, 3, 2)[[0, 1, 3, 2, 4]])
    ry = np.ascontiguousarray(sweep(qy, 2).transpose(0, 1, 3, 2)[[0, 1, 3, 2, 4]])
    qz = np.ascontiguousarray(q.transpose(0, 3, 2, 1)[[0, 1, 4, 3, 2]])
    rz = np.ascontiguousarray(sweep(qz, 3).transpose(0, 3, 2, 1)[[0, 1, 4, 3, 2]])
    # |p|^2 in the floors is summed in physical x,y,z order, so the transposed
    # sweeps agree to rounding rather than bitwise
    assert np.allclose(ry, rx, rtol=1e-13, atol=1e-15)
    assert np.allclose(rz, rx, rtol=1e-13, atol=1e-15)


def test_hll_isothermal_conserves_and_keeps_uniform_state():
    rng = np.random.default_rng(9)
    sh = (1, 6, 20)
    q = cons(rng.uniform(0.5, 2, sh), 0.3 * rng.normal(size=sh), 0.3 * rng.normal(size=sh),
             0.3 * rng.normal(size=sh), rng.uniform(0.5, 2, sh), g=2.0)
    tot0 = q.reshape(5, -1).sum(axis=1)
    po.fluid_update_1d(q, 1, 0.05, 2.0, method=po.HLL, ciso=0.8)
    po.fluid_update_1d(q, 2, 0.05, 2.0, method=po.HLL, ciso=0.8)
    scale = np.abs(q.reshape(5, -1)).sum(axis=1)
    assert np.all(np.abs(q.reshape(5, -1).sum(axis=1) - tot0) <= 1e-13 * scale)
    u = cons(np.full(sh, 1.1), np.full(sh, 0.3), np.full(sh, -0.2), np.zeros(sh), np.full(sh, 0.9), g=2.0)
    r = u.copy()
    po.fluid_update_1d(r, 1, 0.05, 2.0, method=po.HLL, ciso=0.8)
    assert np.allclose(r, u, rtol=1e-15, atol=0)


def _iso_physical_flux(U, pcoef):
    """Mass and momentum fluxes along x with pressure pcoef * rho (pcoef = C_iso^2
    for an isothermal gas)."""
    rho, _, px, py, pz = U
    u = px / rho
    return np.array([px, np.nan, px * u + pcoef * rho, py * u, pz * u])


def _iso_interface_flux(UL, UR, c, pcoefL=None, n=24):
    """The flux the oracle's isothermal HLL sweep puts through the jump of a
    piecewise-constant line (left half UL, right half UR). Next to a jump one
    of the two one-sided differences is zero, so every limiter gives a zero
    slope and the cell left of the jump changes by -lam (F* - F(UL)) + O(lam^2);
    Richardson over lam and lam/2 leaves O(lam^2) ~ 1e-9. pcoefL: the pressure
    coefficient of the flux between two UL cells (C_iso^2 unless that interface
    takes the reference's right-upwind branch)."""
    def rate(lam):
        q = np.zeros((5, 1, 1, n))
        q[:, 0, 0, : n // 2] = UL[:, None]
        q[:, 0, 0, n // 2:] = UR[:, None]
        out = q.copy()
        po.fluid_update_1d(out, 1, lam, 2.0, method=po.HLL, ciso=c)
        return (out - q)[:, 0, 0, n // 2 - 1] / lam
    d = 2 * rate(5e-5) - rate(1e-4)
    return _iso_physical_flux(UL, c * c if pcoefL is None else pcoefL) - d, d


@pytest.mark.parametrize("sign", [1, -1])
def test_hll_isothermal_supersonic_flux_is_the_upwind_physical_flux(sign):
    """Whatever the speed bounds, a jump whose bounds are both of one sign
    passes the physical flux of the upwind state (mass and momenta; the
    energy flux of the isothermal branch is the reference's own). Flow to +x
    takes P = C_iso^2 rho; flow to -x takes the reference's HLL_RIGHT branch,
    whose momentum flux adds C_iso rho, not C_iso^2 rho
    (cudaFluidStep.cu:2191: FLUID_ISOTEMP where the other branches have
    FLUID_ISOTEMPSQR) -- the interface between the two left cells takes that
    branch as well."""
    c = 0.8
    pcoef = c * c if sign > 0 else c
    UL = cons(1.0, 3.0 * sign, 0.2, -0.1, 0.64, g=2.0)
    UR = cons(0.5, 2.8 * sign, -0.3, 0.4, 0.32, g=2.0)
    F, _ = _iso_interface_flux(UL, UR, c, pcoefL=pcoef)
    up = _iso_physical_flux(UL if sign > 0 else UR, pcoef)
    for k in (0, 2, 3, 4):
        assert abs(F[k] - up[k]) <= 1e-7 * max(1.0, abs(up[k])), k


@pytest.mark.parametrize("case", [
    # c, (rho, vx) left, (rho, vx) right
    (0.8, (1.0, 0.2), (0.5, -0.1)),
    (0.8, (1.3, 0.1), (0.4, 0.6)),
    (0.5, (1.0, 0.0), (1.0, 0.6)),   # the -x bound comes from vL - (Vroe - C^2)
    (0.5, (1.2, -0.1), (0.9, 0.55)),
])
def test_hll_isothermal_subsonic_flux_has_the_reference_speed_bounds(case):
    """A subsonic jump: the flux is (SR F_L - SL F_R + SL SR (U_R - U_L)) / (SR - SL)
    for one pair (SL, SR) -- solved from the mass and normal-momentum fluxes, it
    must reproduce both transverse momentum fluxes -- and that pair is the
    reference's computeEinfeldtBounds_isothermal (cudaFluidStep.cu:1779-1799)
    called with C_iso^2 as the sound speed: cs is overwritten by Vroe - cs
    before the one-sided bounds, so SL = min(Vroe - C^2, vL - (Vroe - C^2)) and
    SR = max(Vroe + C^2, vR + (Vroe - C^2)), Vroe the Roe-averaged velocity. The
    last two cases separate this from the textbook Einfeldt bounds
    (vL - C^2, vR + C^2), which differ there by 0.2-0.3. The energy flux is
    then checked against the reference's form with the fitted pair."""
    from scipy.optimize import fsolve
    c, (rl, vl), (rr, vr) = case
    UL = cons(rl, vl, 0.2, -0.1, c * c * rl, g=2.0)
    UR = cons(rr, vr, -0.3, 0.4, c * c * rr, g=2.0)
    F, d = _iso_interface_flux(UL, UR, c)
    FL, FR = _iso_physical_flux(UL, c * c), _iso_physical_flux(UR, c * c)

    def hll(sl, sr):
        return (sr * FL - sl * FR + sl * sr * (UR - UL)) / (sr - sl)

    sl, sr = fsolve(lambda x: [hll(*x)[0] - F[0], hll(*x)[2] - F[2]], [-1.0, 1.0], xtol=1e-14)
    assert np.allclose(hll(sl, sr)[3:], F[3:], rtol=0, atol=1e-8)
    vroe = (np.sqrt(rl) * vl + np.sqrt(rr) * vr) / (np.sqrt(rl) + np.sqrt(rr))
    s = vroe - c * c
    assert abs(sl - min(s, vl - s)) <= 1e-7 and abs(sr - max(vroe + c * c, vr + s)) <= 1e-7
    assert sl < 0 < sr

    # energy, a passive scalar here: the reference's HLL energy flux
    # SR (vL - SL) wL - SL (vR - SR) wR with w = KE + C_iso^2 rho has no
    # 1/(SR - SL) (cudaFluidStep.cu:2178-2181), so between two UL cells it is
    # vL wL (SR - SL) of their own bounds, not vL wL
    def w(U):
        return 0.5 * (U[2] ** 2 + U[3] ** 2 + U[4] ** 2) / U[0] + c * c * U[0]
    sll, srl = min(vl - c * c, c * c), max(vl + c * c, 2 * vl - c * c)
    Fe = vl * w(UL) * (srl - sll) - d[1]
    want = sr * (vl - sl) * w(UL) - sl * (vr - sr) * w(UR)
    assert abs(Fe - want) <= 1e-8 * max(1.0, abs(want))


# ---------------------------------------------------------------- Xin-Jin
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_xj_predictor_is_rusanov_half_step(seed):
    """U = Q - lam/4 [(F+cQ)_i - (F+cQ)_{i-1} - (cQ-F)_{i+1} + (cQ-F)_i]
    = Q - lam/2 (G_{i+1/2} - G_{i-1/2}) with the Rusanov flux G and c the
    line's largest |u| + c_s."""
    n = 20
    U = _line(seed, n, vspread=1.0)
    q = U.reshape(5, 1, 1, n).copy()
    po.xj_predict_1d(q, 1, 0.2, G, rhomin=1e-12)
    rho, v, p = tb.prim(U, G)
    c = np.max(np.abs(v[0]) + np.sqrt(G * p / rho))
    ref = tb.rusanov_step(U, 0.1, G, c)
    err = np.abs(q.reshape(5, n) - ref).max() / np.abs(ref).max()
    assert err < 1e-14, err


def test_xj_floors_closed_form():
    rhomin = 1e-3
    sh = (2, 3, 8)
    q = cons(np.full(sh, 2e-4), np.full(sh, 0.5), np.zeros(sh), np.zeros(sh), np.full(sh, 1e-12))
    r = q.copy()
    po.fluid_update_1d(r, 1, 0.1, G, method=po.XINJIN, rhomin=rhomin)
    # density raised to the floor; momentum and energy are not rescaled (the
    # kinetic energy only drops, so the energy floor stays inactive)
    assert np.all(r[0] == rhomin)
    assert np.allclose(r[1:], q[1:], rtol=1e-14, atol=0)
    # internal energy below rho*emin is raised to it
    rho, vx = 0.7, 1.3
    q = cons(np.full(sh, rho), np.full(sh, vx), np.zeros(sh), np.zeros(sh), np.full(sh, 1e-12))
    r = q.copy()
    po.fluid_update_1d(r, 1, 0.1, G, method=po.XINJIN, rhomin=1e-8)
    emin = 1e-5 / (G - 1)
    assert np.allclose(r[1], 0.5 * rho * vx * vx + rho * emin, rtol=1e-14, atol=0)
    assert np.allclose(r[[0, 2, 3, 4]], q[[0, 2, 3, 4]], rtol=1e-14, atol=0)


def test_xj_predictor_direction_maps_with_symmetric_momentum():
    """The freezing speed always uses |p_x| (cudaFluidStep.cu:2887-2896), so a
    transposed predictor agrees with the X predictor bitwise when
    p_x == p_y == p_z cell by cell. (After the predictor the momenta differ, so
    the full sweeps are not transposes of each other -- that asymmetry is the
    reference's.)"""
    rng = np.random.default_rng(6)
    sh = (9, 9, 9)
    m = rng.normal(size=sh)
    q = cons(rng.uniform(0.5, 2, sh), m, m, m, rng.uniform(1, 2, sh))

    def pred(a, d):
        a = a.copy()
        po.xj_predict_1d(a, d, 0.05, G)
        return a

    rx = pred(q, 1)
    qy = np.ascontiguousarray(q.transpose(0, 1, 3, 2)[[0, 1, 3, 2, 4]])
    ry = np.ascontiguousarray(pred(qy, 2).transpose(0, 1, 3, 2)[[0, 1, 3, 2, 4]])
    qz = np.ascontiguousarray(q.transpose(0, 3, 2, 1)[[0, 1, 4, 3, 2]])
    rz = np.ascontiguousarray(pred(qz, 3).transpose(0, 3, 2, 1)[[0, 1, 4, 3, 2]])
    assert np.array_equal(ry, rx)
    assert np.array_equal(rz, rx)


@pytest.mark.parametrize("method", [po.HLL, po.XINJIN])
def test_uniform_state_is_fixed_point(method):
    sh = (6, 7, 9)
    q = cons(np.full(sh, 1.3), np.full(sh, 0.4), np.full(sh, -0.2), np.full(sh, 0.1), np.full(sh, 0.9))
    r = q.copy()
    for it in range(2):
        po.fluidstep(r, [0.1, 0.1, 0.1], 0.01, 1, it, G, method)
        po.fluidstep(r, [0.1, 0.1, 0.1], 0.01, -1, it, G, method)
    assert np.allclose(r, q, rtol=2e-15, atol=0)


@pytest.mark.parametrize("method", [po.HLL, po.XINJIN])
def test_periodic_conservation(method):
    rng = np.random.default_rng(3)
    sh = (10, 12, 14)
    q = cons(rng.uniform(0.5, 2, sh), rng.normal(size=sh) * 0.3, rng.normal(size=sh) * 0.3,
             rng.normal(size=sh) * 0.3, rng.uniform(0.5, 2, sh))
    tot0 = q.reshape(5, -1).sum(axis=1)
    for it in range(3):
        po.fluidstep(q, [0.1, 0.1, 0.1], 0.004, 1, it, G, method)
        po.fluidstep(q, [0.1, 0.1, 0.1], 0.004, -1, it, G, method)
    scale = np.abs(q.reshape(5, -1)).sum(axis=1)
    assert np.all(np.abs(q.reshape(5, -1).sum(axis=1) - tot0) <= 1e-13 * scale)


def _run_1d(q, h, tend, method, cfl, modes):
    t, it = 0.0, 0
    N = q.shape[-1]
    while t < tend * (1 - 1e-14):
        cs = po.soundspeed(q, G)
        tau = po.cfl_timestep(q, cs, [h, 1, 1], method, (N, 1, 1))
        dt = cfl * tau
        if t + 2 * dt > tend:
            dt = 0.5 * (tend - t)
        po.fluidstep(q, [h, 1, 1], dt, 1, it, G, method, po.LIM_VANLEER, modes=modes)
        po.fluidstep(q, [h, 1, 1], dt, -1, it, G, method, po.LIM_VANLEER, modes=modes)
        t += 2 * dt
        it += 1
    return q


@pytest.mark.parametrize("method,cfl", [(po.HLL, 0.4), (po.XINJIN, 0.4)])
def test_sod_converges(method, cfl):
    errs = []
    for N in (100, 200, 400):
        dx = 1.0 / N
        x = (np.arange(N) + 0.5) * dx
        q = cons(np.where(x < 0.5, 1.0, 0.125), 0 * x, 0 * x, 0 * x, np.where(x < 0.5, 1.0, 0.1))
        q = q.reshape(5, 1, 1, N).copy()
        _run_1d(q, dx, 0.2, method, cfl, (po.BC_CONST,) * 6)
        ex = er.cell_averages(x, dx, 0.5, 0.2, (1, 0, 1), (0.125, 0, 0.1), G)
        errs.append(np.abs(q[0, 0, 0] - ex[0]).mean())
    assert errs[0] < 8e-3 and errs[2] < 3e-3, errs
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 0.7), (errs, orders)


@pytest.mark.parametrize("method", [po.HLL, po.XINJIN])
def test_smooth_wave_second_order(method):
    errs = []
    for N in (32, 64, 128):
        dx = 1.0 / N
        x = (np.arange(N) + 0.5) * dx

        def avg(t):
            return 1 + 0.2 * (np.cos(2 * np.pi * (x - dx / 2 - t)) -
                              np.cos(2 * np.pi * (x + dx / 2 - t))) / (2 * np.pi * dx)

        q = cons(avg(0.0), np.ones(N), 0 * x, 0 * x, np.ones(N)).reshape(5, 1, 1, N).copy()
        _run_1d(q, dx, 0.25, method, 0.4, (po.BC_CIRCULAR,) * 6)
        errs.append(np.abs(q[0, 0, 0] - avg(0.25)).mean())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert orders[-1] > 1.6, (errs, orders)
