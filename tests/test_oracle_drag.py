"""Pins the oracle's gas-dust drag (oracle/drag_oracle.c).

What pins what:
  drag constants + rate k   the independent dusty-box formulation of the
                            acceleration (experiment/Advection/DustyBoxAnalyzer.m:
                            91-104, 152-155), written out here from its text
  exact integrator          scipy's DOP853 at rtol 1e-13 on that same ODE in
                            |dv| (not ln|dv|), with P(|dv|) from energy
                            conservation (DustyBoxAnalyzer.m:135-140)
  LogTrap2 / LogTrap3       their convergence orders 2 and 3 against the exact
                            solution (the reference's tsDustybox.m criterion)
  ETDRK1                    exact in the linear (Stokes, no heating) limit;
                            local truncation order 2 in general
  EMP / RK4 on |dv|         local truncation orders 3 and 5 (global 2 and 4)
                            against the exact solution; in the linear-drag
                            limit the Taylor polynomials of exp(-x) to 2nd
                            and 4th order, which fix every stage weight
  the cell update           exact momentum and total-energy conservation,
                            unchanged barycentric velocity, dv direction kept
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import pyoracle as po

AMU = 1.66053904e-27
KB = 1.381e-23


def gas_dust_thermo():
    """fluidDetailModel('cold_molecular_hydrogen') and ('10um_iron_balls') with
    the dust scaled as run/profile_3DXYZ_gasdust.m:43-46 does (sigma/10, m/1e4)."""
    m = 2.016 * AMU
    mu0 = 8.9135e-6
    sigma = np.sqrt(2 / 27) * 1.7 * np.sqrt(KB * 298.15 * m) / mu0  # generateSigma
    tg = po.Thermo(5 / 3, m, -1.0, KB, mu0, 0.7, sigma, 0.2)
    td = po.Thermo(1.01, 3e-11 * 1e-4, 0.1, KB, 0.0, 0.0, np.pi * 25e-12 * 1e-1, 0.0)
    return tg, td


def dustybox_accel(tg, td, rhog, rhod, P, dv):
    """d|dv|/dt of DustyBoxAnalyzer.computeAcceleration, as written there."""
    T = (tg.m * P) / (rhog * tg.kBolt)
    visc = tg.mu0 * (T / 298.15) ** tg.muTindex
    Rey = rhog * dv * np.sqrt(td.sigma0 / np.pi) / visc
    mfp = tg.m * (T / 298.15) ** tg.sigmaTindex / (rhog * tg.sigma0 * np.sqrt(2))
    Kn = 2 * mfp / np.sqrt(td.sigma0 / np.pi)
    C = (24 / Rey + 3.6 * Rey ** -.319 + .4072 * Rey / (8710 + Rey)) / (1 + Kn * (1.142 + 0.558 * np.exp(-0.999 / Kn)))
    return -C * .5 * dv ** 2 * (td.sigma0 / 4) * (rhog + rhod) / td.m


def test_drag_rate_matches_dustybox_formulation():
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rng = np.random.default_rng(0)
    for _ in range(200):
        rhog = rng.uniform(0.1, 3)
        rhod = rng.uniform(0.01, 3)
        P = rng.uniform(0.05, 5)
        dv = 10 ** rng.uniform(-6, 0.5)
        u = P / ((tg.gamma - 1) * rhog)
        k = po.drag_k(P9, rhog, rhod, dv, u)
        a = dustybox_accel(tg, td, rhog, rhod, P, dv)
        assert k * dv == pytest.approx(-a, rel=1e-12)
    assert po.drag_k(P9, 1.0, 0.1, 0.9e-9, 1.0) == 0.0  # no drag below |dv| = 1e-9 (:853-859)


@pytest.mark.parametrize("case", range(6))
def test_exact_integrator_matches_dop853(case):
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rng = np.random.default_rng(10 + case)
    rhog = rng.uniform(0.3, 2)
    rhod = rng.uniform(0.02, 2)
    P0 = rng.uniform(0.2, 3)
    dv0 = 10 ** rng.uniform(-3, 0)
    k0 = po.drag_k(P9, rhog, rhod, dv0, P0 / ((tg.gamma - 1) * rhog))
    t = (0.02, 0.2, 1.0, 3.0, 10.0, 0.5)[case] / k0  # from mild to very stiff
    mu = rhog * rhod / (rhog + rhod)

    def rhs(_t, y):
        P = P0 + (tg.gamma - 1) * .5 * mu * (dv0 ** 2 - y[0] ** 2)
        return [dustybox_accel(tg, td, rhog, rhod, P, y[0])]

    sol = solve_ivp(rhs, (0, t), [dv0], method="DOP853", rtol=1e-13, atol=1e-20)
    ref = sol.y[0, -1] / dv0
    r, steps = po.drag_exact_ratio(P9, rhog, rhod, dv0, P0 / ((tg.gamma - 1) * rhog), t, 1e-12)
    assert steps > 0
    assert r == pytest.approx(ref, rel=2e-11), (r, ref, steps)


def one_cell(rhog, vg, Pg, rhod, vd, gamma=5 / 3):
    g = np.zeros((5, 1, 1, 1))
    d = np.zeros((5, 1, 1, 1))
    g[0] = rhog
    g[2:, 0, 0, 0] = rhog * np.asarray(vg)
    g[1] = Pg / (gamma - 1) + .5 * rhog * np.sum(np.square(vg))
    d[0] = rhod
    d[2:, 0, 0, 0] = rhod * np.asarray(vd)
    d[1] = .5 * rhod * np.sum(np.square(vd))
    return g, d


def rel_dv(g, d):
    return g[2:, 0, 0, 0] / g[0, 0, 0, 0] - d[2:, 0, 0, 0] / d[0, 0, 0, 0]


def test_logtrap_convergence_orders():
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rhog, rhod, Pg = 1.0, 0.5, 1.0
    vg, vd = [0.4, -0.1, 0.2], [-0.3, 0.1, 0.0]
    g0, d0 = one_cell(rhog, vg, Pg, rhod, vd)
    s0 = np.linalg.norm(rel_dv(g0, d0))
    k0 = po.drag_k(P9, rhog, rhod, s0, Pg / ((tg.gamma - 1) * rhog))
    T = 2.0 / k0
    exact, _ = po.drag_exact_ratio(P9, rhog, rhod, s0, Pg / ((tg.gamma - 1) * rhog), T, 1e-13)
    for method, order in ((po.DRAG_LOGTRAP2, 2), (po.DRAG_LOGTRAP3, 3)):
        errs = []
        for n in (8, 16, 32, 64):
            g, d = g0.copy(), d0.copy()
            for _ in range(n):
                po.drag(g, d, tg, td, T / n, method)
            errs.append(abs(np.linalg.norm(rel_dv(g, d)) / s0 - exact))
        orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
        assert np.all(orders > order - 0.3), (method, errs, orders)
        assert np.all(orders < order + 0.5), (method, errs, orders)


@pytest.mark.parametrize("method", [po.DRAG_LOGTRAP2, po.DRAG_LOGTRAP3, po.DRAG_EXACT, po.DRAG_ETDRK1])
def test_drag_invariants(method):
    tg, td = gas_dust_thermo()
    rng = np.random.default_rng(4)
    sh = (4, 5, 6)
    g = np.zeros((5,) + sh)
    d = np.zeros((5,) + sh)
    g[0] = rng.uniform(0.5, 2, sh)
    d[0] = g[0] * np.exp(rng.normal(np.log(0.1), 1.0, sh))
    vg = rng.normal(0, 0.4, (3,) + sh)
    vd = rng.normal(0, 0.4, (3,) + sh)
    vd[:, 0, 0, :2] = vg[:, 0, 0, :2]  # two cells without relative motion
    g[2:] = g[0] * vg
    d[2:] = d[0] * vd
    g[1] = rng.uniform(0.5, 2, sh) / (tg.gamma - 1) + .5 * g[0] * (vg ** 2).sum(0)
    d[1] = .5 * d[0] * (vd ** 2).sum(0)
    g0, d0 = g.copy(), d.copy()
    dv0 = g0[2:] / g0[0] - d0[2:] / d0[0]
    po.drag(g, d, tg, td, 0.02, method)
    if method == po.DRAG_ETDRK1:
        # the reference's driven form 0/k + (dv - 0/k) e^{-kt} is NaN where the
        # rate is cut to 0 (|dv| < 1e-9); compare the moving cells only
        assert np.all(np.isnan(g[2:, 0, 0, :2])) and np.all(np.isnan(d[2:, 0, 0, :2]))
        for a, b in ((g, g0), (d, d0)):
            a[:, 0, 0, :2] = b[:, 0, 0, :2]
    # momentum per cell
    assert np.allclose(g[2:] + d[2:], g0[2:] + d0[2:], rtol=0, atol=1e-14)
    # total energy per cell: gas E + dust kinetic energy
    ekd0 = .5 * (d0[2:] ** 2).sum(0) / d0[0]
    ekd = .5 * (d[2:] ** 2).sum(0) / d[0]
    assert np.allclose(g[1] + ekd, g0[1] + ekd0, rtol=1e-14, atol=1e-14)
    # density untouched; dust energy slab untouched
    assert np.array_equal(g[0], g0[0]) and np.array_equal(d[0], d0[0]) and np.array_equal(d[1], d0[1])
    # dv keeps its direction and shrinks
    dv = g[2:] / g[0] - d[2:] / d[0]
    n0 = np.sqrt((dv0 ** 2).sum(0))
    n1 = np.sqrt((dv ** 2).sum(0))
    moving = n0 > 1e-9
    assert np.all(n1[moving] < n0[moving])
    cos = (dv * dv0).sum(0)[moving] / (n0 * n1)[moving]
    assert np.all(cos > 1 - 1e-12)
    assert np.allclose(dv[:, 0, 0, :2], 0, atol=1e-15)


def test_emp_rk4_local_orders_against_exact():
    """One step of size dt: the |dv| error shrinks as dt^3 (EMP) and dt^5 (RK4)."""
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rhog, rhod, Pg = 1.0, 0.3, 1.0
    u0 = Pg / ((tg.gamma - 1) * rhog)
    s0 = 0.5
    k0 = po.drag_k(P9, rhog, rhod, s0, u0)
    for method, local, xs in ((po.DRAG_EMP, 3, (0.4, 0.2, 0.1, 0.05)), (po.DRAG_RK4, 5, (0.1, 0.05, 0.025))):
        errs = []
        for x in xs:
            dt = x / k0
            exact, _ = po.drag_exact_ratio(P9, rhog, rhod, s0, u0, dt, 1e-14)
            errs.append(abs(po.drag_rk_magdv(P9, rhog, rhod, s0, u0, dt, method) - exact * s0))
        orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
        assert np.all(orders > local - 0.45), (method, errs, orders)
        assert np.all(orders < local + 0.6), (method, errs, orders)


def test_emp_rk4_linear_limit_is_taylor_polynomial():
    """With k independent of |dv| and no heating (Stokes regime, Re -> 0,
    Kn -> 0, rho_d -> 0): d|dv|/dt = -K |dv|, and one step must return
    s0 (1 - x + x^2/2) for EMP and s0 sum_{j<=4} (-x)^j / j! for RK4, x = K dt.
    A wrong stage node or weight changes a coefficient."""
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rhog, rhod, u0 = 100.0, 0.5, 1.0  # Kn ~ 1e-10
    s0 = 1e-21                        # Re ~ 2e-20: the .45 Re^-.319 term is 1e-14 of 3/Re
    # Stokes-limit rate from the constants alone: Cd = (3/Re) / Cunningham,
    # k = Cd |dv| (rho_g + rho_d) epsilon, independent of |dv|
    T = P9[3] * u0
    Kn = P9[4] * T ** P9[2] / rhog
    cunningham = 1 + Kn * (1.142 + 0.558 * np.exp(-0.999 / Kn))
    K = 3 * (rhog + rhod) * P9[6] / (P9[5] * rhog * T ** P9[1] * cunningham)
    for method, coeffs in ((po.DRAG_EMP, [1, -1, 1 / 2]), (po.DRAG_RK4, [1, -1, 1 / 2, -1 / 6, 1 / 24])):
        for x in (0.3, 0.6, 0.9):  # larger x drives an RK4 stage |dv| negative (NaN, as in the reference)
            got = po.drag_rk_magdv(P9, rhog, rhod, s0, u0, x / K, method) / s0
            want = sum(c * x ** j for j, c in enumerate(coeffs))
            assert got == pytest.approx(want, rel=2e-8), (method, x, got, want)


@pytest.mark.parametrize("method", [po.DRAG_EMP, po.DRAG_RK4])
def test_emp_rk4_cell_update(method):
    """cukern_applyFinalDeltaV: momentum conserved, gas E + dust KE conserved,
    dv keeps its direction with |dv| = the integrator's value, and the dust
    energy reset to its kinetic energy + 1e-4 rho_d."""
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rng = np.random.default_rng(9)
    sh = (3, 4, 5)
    g = np.zeros((5,) + sh)
    d = np.zeros((5,) + sh)
    g[0] = rng.uniform(0.5, 2, sh)
    d[0] = g[0] * np.exp(rng.normal(np.log(0.1), 1.0, sh))
    vg = rng.normal(0, 0.4, (3,) + sh)
    vd = rng.normal(0, 0.4, (3,) + sh)
    g[2:] = g[0] * vg
    d[2:] = d[0] * vd
    g[1] = rng.uniform(0.5, 2, sh) / (tg.gamma - 1) + .5 * g[0] * (vg ** 2).sum(0)
    d[1] = .5 * d[0] * (vd ** 2).sum(0) + 0.3
    g0, d0 = g.copy(), d.copy()
    # explicit methods are conditionally stable: keep k dt <= 1/2 in every cell
    # (beyond it a stage |dv| can go negative and the reference gives NaN)
    kmax = 0.0
    for idx in np.ndindex(*sh):
        c = (slice(None),) + idx
        s0 = np.linalg.norm(g0[2:][c] / g0[0][idx] - d0[2:][c] / d0[0][idx])
        u0 = (g0[1][idx] - .5 * (g0[2:][c] ** 2).sum() / g0[0][idx]) / g0[0][idx]
        kmax = max(kmax, po.drag_k(P9, g0[0][idx], d0[0][idx], s0, u0))
    dt = 0.5 / kmax
    po.drag(g, d, tg, td, dt, method)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(d))
    assert np.allclose(g[2:] + d[2:], g0[2:] + d0[2:], rtol=0, atol=1e-14)
    ekd0 = .5 * (d0[2:] ** 2).sum(0) / d0[0]
    ekd = .5 * (d[2:] ** 2).sum(0) / d[0]
    assert np.allclose(g[1] + ekd, g0[1] + ekd0, rtol=1e-14, atol=1e-14)
    assert np.allclose(d[1], ekd + 1e-4 * d[0], rtol=1e-15, atol=0)
    dv0 = g0[2:] / g0[0] - d0[2:] / d0[0]
    dv = g[2:] / g[0] - d[2:] / d[0]
    n0 = np.sqrt((dv0 ** 2).sum(0))
    n1 = np.sqrt((dv ** 2).sum(0))
    assert np.all((dv * dv0).sum(0) / (n0 * n1) > 1 - 1e-12)
    for idx in np.ndindex(*sh):
        u0 = (g0[1][idx] - .5 * (g0[2:][(slice(None),) + idx] ** 2).sum() / g0[0][idx]) / g0[0][idx]
        want = po.drag_rk_magdv(P9, g0[0][idx], d0[0][idx], n0[idx], u0, dt, method)
        assert n1[idx] == pytest.approx(want, rel=1e-12)


def test_etdrk1_decay_and_local_order():
    """Exponential Euler (solveDragETDRK1): one step scales |dv| by exactly
    exp(-k0 dt), k0 the rate of the initial state (so it is exact whenever the
    rate is constant along the path); against the exact relaxation its
    one-step error shrinks as dt^2 (a wrong rate state or a dropped heating
    term would not)."""
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    rhog, rhod, Pg = 1.0, 0.4, 1.0
    g0, d0 = one_cell(rhog, [0.4, -0.1, 0.2], Pg, rhod, [-0.3, 0.1, 0.0])
    s0 = np.linalg.norm(rel_dv(g0, d0))
    u0 = Pg / ((tg.gamma - 1) * rhog)
    k0 = po.drag_k(P9, rhog, rhod, s0, u0)
    for x in (0.3, 2.0, 9.0):
        g, d = g0.copy(), d0.copy()
        po.drag(g, d, tg, td, x / k0, po.DRAG_ETDRK1)
        assert np.linalg.norm(rel_dv(g, d)) == pytest.approx(s0 * np.exp(-x), rel=1e-12)
    errs = []
    for x in (0.2, 0.1, 0.05, 0.025):
        exact, _ = po.drag_exact_ratio(P9, rhog, rhod, s0, u0, x / k0, 1e-14)
        g, d = g0.copy(), d0.copy()
        po.drag(g, d, tg, td, x / k0, po.DRAG_ETDRK1)
        errs.append(abs(np.linalg.norm(rel_dv(g, d)) / s0 - exact))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 1.7) and np.all(orders < 2.4), (errs, orders)


# ---------------------------------------------------------------- invalid states
@pytest.mark.parametrize("cell", [
    dict(rhog=float("nan"), rhod=0.1, s0=0.2, u0=1.0),
    dict(rhog=1.0, rhod=0.1, s0=float("nan"), u0=1.0),
    dict(rhog=1.0, rhod=0.1, s0=float("inf"), u0=1.0),
    dict(rhog=1.0, rhod=0.1, s0=0.2, u0=float("nan")),
    dict(rhog=1.0, rhod=0.1, s0=0.2, u0=-0.5),
    dict(rhog=1.0, rhod=0.1, s0=0.2, u0=0.0),
    dict(rhog=-1.0, rhod=0.1, s0=0.2, u0=1.0),
    dict(rhog=1.0, rhod=0.0, s0=0.2, u0=1.0),
])
def test_exact_integrator_returns_at_once_on_invalid_cells(cell):
    """A cell the relaxation problem is not defined on returns NaN before any
    step, flagged invalid -- not a spin to the step budget (a NaN error
    estimate used to fail `err <= 1` and grow the step forever)."""
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    r, steps = po.drag_exact_ratio(P9, cell["rhog"], cell["rhod"], cell["s0"], cell["u0"], 1.0)
    assert np.isnan(r)
    assert steps == po.DRAG_INVALID


def test_exact_integrator_budget_overrun_is_reported():
    """A tolerance no step size can meet exhausts the 4096-step budget and
    says so with a negative count."""
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    u0 = 1.0 / ((tg.gamma - 1) * 1.0)
    r, steps = po.drag_exact_ratio(P9, 1.0, 0.5, 0.3, u0, 1.0, tol=1e-40)
    assert steps == -4096
    assert np.isfinite(r)


def test_drag_call_reports_invalid_cells_and_still_updates_the_rest():
    tg, td = gas_dust_thermo()
    rng = np.random.default_rng(41)
    sh = (4, 6, 50)
    g = np.zeros((5,) + sh)
    d = np.zeros((5,) + sh)
    g[0] = rng.uniform(0.5, 2, sh)
    d[0] = rng.uniform(0.05, 0.5, sh)
    for c in range(2, 5):
        g[c] = g[0] * rng.normal(0, 0.3, sh)
        d[c] = d[0] * rng.normal(0, 0.3, sh)
    g[1] = 1.0 / (tg.gamma - 1) + 0.5 * (g[2] ** 2 + g[3] ** 2 + g[4] ** 2) / g[0]
    d[1] = 0.5 * (d[2] ** 2 + d[3] ** 2 + d[4] ** 2) / d[0]
    bad_g, bad_d = g.copy(), d.copy()
    bad_g[:, 1, 2, 3] = np.nan                 # a NaN cell
    bad_g[1, 2, 4, 10] = 0.0                   # negative internal energy
    bad_d[0, 3, 5, 49] = -1.0                  # negative dust density
    ok_g, ok_d = g.copy(), d.copy()
    po.drag(ok_g, ok_d, tg, td, 0.05, method=po.DRAG_EXACT)
    with pytest.raises(po.DragStateError):
        po.drag(bad_g, bad_d, tg, td, 0.05, method=po.DRAG_EXACT)
    mask = np.ones(sh, bool)
    mask[1, 2, 3] = mask[2, 4, 10] = mask[3, 5, 49] = False
    assert np.array_equal(bad_g[:, mask], ok_g[:, mask])
    assert np.array_equal(bad_d[:, mask], ok_d[:, mask])
    assert np.isnan(bad_g[2, 2, 4, 10]) and np.isnan(bad_d[2, 3, 5, 49])
