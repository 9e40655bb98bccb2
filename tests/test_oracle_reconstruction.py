"""Pins the oracle's second-order reconstruction (oracle/imogen_oracle.c)
against independent textbook formulations (tests/hydro_textbook.py):

  HLLC sweep   MUSCL on (rho, internal energy density, velocity) with the
               limiter's r-form, textbook HLLC (Toro 10.37-10.39) on the
               faces, SSP-RK2 -- the choice of cudaFluidStep.cu:1218-1311
               (RECONSTRUCT_VELOCITY, face P = (gamma-1)(e -+ s)). The same
               scheme reconstructing momentum, total energy or the specific
               internal energy instead differs by far more than the tolerance.
  HLL sweep    MUSCL on (rho, velocity, pressure) with textbook HLLE
               (cudaFluidStep.cu:1887-1889, 2124-2196)
  Xin-Jin      the relaxed Jin-Xin scheme in interface-flux form, with
               MUSCL-corrected characteristic variables c U -+ F in the
               corrector (cudaFluidStep.cu:2541-2565) and a per-line frozen
               speed (:2868-2941)
  limiters     superbee's and van Albada's closed forms (limiter_functions.h:
               80-106), including superbee's missing 1/2 on its first branch
               and van Albada's 0/0 at A = B = 0
  smoothness   self-convergence at ~2nd order on a nonlinear acoustic wave,
               where velocity and pressure vary (HLLC, HLL, Xin-Jin)

Internal energy density and pressure differ by the constant factor gamma-1
and every limiter here is homogeneous of degree one, so reconstructing P
instead of e is the same scheme; the tests separate the variable choices
that are not.
"""
import numpy as np
import pytest

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


def smooth_line(seed, n):
    """Random but resolved periodic line: a few Fourier modes per variable, so
    the limiter acts in every regime (smooth slopes, extrema, sign changes)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n

    def field(base, amp):
        f = np.full(n, base)
        for k in range(1, 4):
            f += amp * rng.uniform(-1, 1) * np.sin(2 * np.pi * k * x + rng.uniform(0, 2 * np.pi)) / k
        return f

    return cons(field(1.0, 0.3), field(0.0, 0.6), field(0.0, 0.4), field(0.0, 0.4), field(1.0, 0.3))


def rough_line(seed, n):
    rng = np.random.default_rng(seed)
    return cons(rng.uniform(0.3, 2.0, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                rng.uniform(0.3, 2.0, n))


def oracle_sweep(U, lam, method, lim):
    n = U.shape[1]
    q = U.reshape(5, 1, 1, n).copy()
    po.fluid_update_1d(q, 1, lam, G, method=method, limiter=lim, rhomin=1e-12)
    return q.reshape(5, n)


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def superbee_half(A, B):
    """The reference's superbee half-slope by cases: same signs, A, B > 0:
    min(A, 2B) when B < A (the full slope: no 1/2 on this branch) else
    min(A, B/2); A, B < 0: A when |B| > |A| else B/2; opposite signs: 0."""
    A = np.asarray(A, dtype=float)
    B = np.asarray(B, dtype=float)
    pos = np.where(B < A, np.minimum(A, 2 * B), np.minimum(A, 0.5 * B))
    neg = np.where(np.abs(B) > np.abs(A), A, 0.5 * B)
    return np.where(A * B < 0, 0.0, np.where(A > 0, pos, np.where(A < 0, neg, 0.0)))


SLOPES = {
    po.LIM_OSPRE: lambda A, B: tb.half_slope(tb.psi_ospre, A, B),
    po.LIM_VANLEER: lambda A, B: tb.half_slope(tb.psi_vanleer, A, B),
    po.LIM_MC: lambda A, B: tb.half_slope(tb.psi_mc, A, B),
    po.LIM_VANALBADA: lambda A, B: tb.half_slope(tb.psi_vanalbada, A, B),
    po.LIM_SUPERBEE: superbee_half,
}


# ---------------------------------------------------------------- limiter closed forms
def test_superbee_closed_form():
    rng = np.random.default_rng(21)
    pairs = list(rng.normal(size=(400, 2))) + [(1.0, 0.3), (0.3, 1.0), (-1.0, -0.3), (-0.3, -1.0), (2.0, 1.5),
                                               (1.5, 2.0), (0.0, 1.0), (1.0, 0.0), (-1.0, 0.0), (0.0, -1.0)]
    for a, b in pairs:
        assert po.slope_limiter(po.LIM_SUPERBEE, a, b) == pytest.approx(float(superbee_half(a, b)), rel=1e-15,
                                                                           abs=1e-300)
    # the missing 1/2: on B < A (both positive) it returns the full textbook
    # superbee slope max(min(A, 2B), min(2A, B)), on B >= A half of it
    for a, b in [(1.0, 0.3), (1.0, 0.7), (0.3, 1.0), (0.7, 1.0)]:
        full = max(min(a, 2 * b), min(2 * a, b))
        got = po.slope_limiter(po.LIM_SUPERBEE, a, b)
        assert got == (full if b < a else 0.5 * full)


def test_vanalbada_closed_form():
    rng = np.random.default_rng(22)
    for a, b in rng.normal(size=(400, 2)):
        exp = float(tb.half_slope(tb.psi_vanalbada, a, b))
        assert po.slope_limiter(po.LIM_VANALBADA, a, b) == pytest.approx(exp, rel=1e-14, abs=1e-300)
    # A = B gives the plain half difference; one zero difference gives 0; 0/0 at A = B = 0
    for a in (0.1, 1.0, -3.0):
        assert po.slope_limiter(po.LIM_VANALBADA, a, a) == pytest.approx(0.5 * a, rel=1e-15)
        assert po.slope_limiter(po.LIM_VANALBADA, a, 0.0) == 0.0
        assert po.slope_limiter(po.LIM_VANALBADA, 0.0, a) == 0.0
    assert np.isnan(po.slope_limiter(po.LIM_VANALBADA, 0.0, 0.0))


@pytest.mark.parametrize("lim", [po.LIM_OSPRE, po.LIM_VANLEER, po.LIM_MC, po.LIM_VANALBADA])
def test_limiters_match_their_r_form(lim):
    rng = np.random.default_rng(23)
    for a, b in rng.normal(size=(300, 2)):
        exp = float(SLOPES[lim](a, b))
        assert po.slope_limiter(lim, a, b) == pytest.approx(exp, rel=1e-14, abs=1e-300)


def test_symmetric_limiters_give_half_differences_on_linear_data():
    """On equal backward and forward differences (a linear profile) the
    symmetric limiters return exactly half the difference."""
    for lim in (po.LIM_OSPRE, po.LIM_VANLEER, po.LIM_MC, po.LIM_VANALBADA):
        for d in (0.125, -0.375, 3.0):
            assert po.slope_limiter(lim, d, d) == 0.5 * d


# ---------------------------------------------------------------- HLLC MUSCL
@pytest.mark.parametrize("lim", sorted(SLOPES))
@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("kind", ["smooth", "rough"])
def test_hllc_sweep_is_muscl_on_density_energy_velocity(lim, seed, kind):
    n = 40
    U = smooth_line(seed, n) if kind == "smooth" else rough_line(seed, n)
    lam = 0.08
    got = oracle_sweep(U, lam, po.HLLC, lim)
    ref = tb.muscl_heun_line(tb.hllc_textbook, U, lam, G, SLOPES[lim], recon="prim")
    assert relerr(got, ref) < 1e-13


@pytest.mark.parametrize("recon", ["cons", "prim_p_rho"])
def test_other_reconstruction_variables_are_rejected(recon):
    """The same scheme on other variables is a different scheme: each of these
    plausible substitutions misses the oracle by orders of magnitude more
    than the 1e-13 the right one meets."""
    U = smooth_line(4, 40)
    got = oracle_sweep(U, 0.08, po.HLLC, po.LIM_OSPRE)
    wrong = tb.muscl_heun_line(tb.hllc_textbook, U, 0.08, G, SLOPES[po.LIM_OSPRE], recon=recon)
    assert relerr(got, wrong) > 1e-6


def test_linear_profile_faces_are_the_interface_midpoints():
    """rho, e and v linear over the whole line segment: Ospre and van Leer
    return exactly half the difference, so both faces of every interior
    interface are the midpoint state and the first stage's flux there is
    the physical flux of that state. Reconstructing momentum instead would
    put rho v at the midpoint average, off by drho dv / 4."""
    n = 16
    i = np.arange(n, dtype=float)
    rho = 1.0 + 0.05 * i
    e = 2.0 + 0.1 * i
    v = np.vstack([0.3 - 0.02 * i, 0.1 + 0.01 * i, -0.2 + 0.03 * i])
    U = tb.to_cons(rho, e, v)
    for lim in (po.LIM_OSPRE, po.LIM_VANLEER):
        L, R = tb.muscl_faces(U, G, SLOPES[lim], "prim")
        for k in range(1, n - 2):
            mid = tb.to_cons(0.5 * (rho[k] + rho[k + 1]), 0.5 * (e[k] + e[k + 1]),
                             0.5 * (v[:, k] + v[:, k + 1]))
            assert np.allclose(R[:, k], mid, rtol=1e-15, atol=0)
            assert np.allclose(L[:, k + 1], mid, rtol=1e-15, atol=0)
        # stage-A flux difference of an interior cell from the oracle: a
        # single-stage step is not exposed, so compare the full sweep with the
        # textbook scheme on this line (periodic wrap included)
        got = oracle_sweep(U, 0.05, po.HLLC, lim)
        ref = tb.muscl_heun_line(tb.hllc_textbook, U, 0.05, G, SLOPES[lim], recon="prim")
        assert relerr(got, ref) < 1e-13
        wrong = tb.muscl_heun_line(tb.hllc_textbook, U, 0.05, G, SLOPES[lim], recon="cons")
        assert relerr(got, wrong) > 1e-7


# ---------------------------------------------------------------- HLL MUSCL
def _hll_slope_prim(U, g, slope):
    return tb.muscl_faces(U, g, slope, "prim")


@pytest.mark.parametrize("lim", sorted(SLOPES))
@pytest.mark.parametrize("seed", [0, 1])
def test_hll_sweep_is_muscl_on_density_velocity_pressure(lim, seed):
    """HLL reconstructs (rho, v, P) (cudaFluidStep.cu:1887-1889); P = (gamma-1) e
    with homogeneous limiters makes this the same MUSCL as the HLLC sweep's."""
    n = 40
    U = smooth_line(seed + 10, n)
    lam = 0.08
    got = oracle_sweep(U, lam, po.HLL, lim)
    ref = tb.muscl_heun_line(tb.hlle_textbook, U, lam, G, SLOPES[lim], recon="prim")
    assert relerr(got, ref) < 1e-13
    wrong = tb.muscl_heun_line(tb.hlle_textbook, U, lam, G, SLOPES[lim], recon="cons")
    assert relerr(got, wrong) > 1e-7


# ---------------------------------------------------------------- Xin-Jin
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["smooth", "rough"])
def test_xinjin_sweep_is_relaxed_jin_xin_with_ospre_corrector(seed, kind):
    n = 40
    U = smooth_line(seed + 20, n) if kind == "smooth" else rough_line(seed + 20, n)
    lam = 0.05
    got = oracle_sweep(U, lam, po.XINJIN, po.LIM_OSPRE)
    ref = tb.jin_xin_line(U, lam, G, rhomin=1e-12, emin=1e-5 / (G - 1))
    assert relerr(got, ref) < 1e-13
    # a dropped TVD term (first-order corrector) or a limiter on the
    # conserved variables instead of c U -+ F both miss by far more
    first = tb.jin_xin_line(U, lam, G, psi=lambda r: 0.0 * r, rhomin=1e-12, emin=1e-5 / (G - 1))
    assert relerr(got, first) > 1e-6
    vl = tb.jin_xin_line(U, lam, G, psi=tb.psi_vanleer, rhomin=1e-12, emin=1e-5 / (G - 1))
    assert relerr(got, vl) > 1e-8


# ---------------------------------------------------------------- acoustic convergence
def _acoustic(N, amp=0.05):
    """Right-going simple wave of finite amplitude: Riemann invariant J- and
    entropy uniform, so u and P vary with rho (steepens, but stays smooth
    before t = 0.1 at this amplitude)."""
    x = (np.arange(N) + 0.5) / N
    rho0, P0 = 1.0, 1.0 / G
    c0 = np.sqrt(G * P0 / rho0)
    rho = rho0 * (1 + amp * np.sin(2 * np.pi * x))
    P = P0 * (rho / rho0) ** G
    c = np.sqrt(G * P / rho)
    u = 2 / (G - 1) * (c - c0)
    return cons(rho, u, 0 * x, 0 * x, P).reshape(5, 1, 1, N).copy()


def _evolve(q, N, tend, method, lim):
    h = 1.0 / N
    t, it = 0.0, 0
    while t < tend * (1 - 1e-14):
        cs = po.soundspeed(q, G)
        dt = 0.4 * po.cfl_timestep(q, cs, [h, 1, 1], method, (N, 1, 1))
        if t + 2 * dt > tend:
            dt = 0.5 * (tend - t)
        po.fluidstep(q, [h, 1, 1], dt, 1, it, G, method, lim, modes=(po.BC_CIRCULAR,) * 6)
        po.fluidstep(q, [h, 1, 1], dt, -1, it, G, method, lim, modes=(po.BC_CIRCULAR,) * 6)
        t += 2 * dt
        it += 1
    return q


@pytest.mark.parametrize("method,lim", [(po.HLLC, po.LIM_VANLEER), (po.HLLC, po.LIM_OSPRE),
                                        (po.HLL, po.LIM_OSPRE), (po.XINJIN, po.LIM_OSPRE)])
def test_acoustic_wave_self_convergence_second_order(method, lim):
    """Errors between successive resolutions (fine solution averaged onto the
    coarse grid) fall at ~2nd order in L1 on every conserved variable: a
    momentum reconstruction would still converge, but a dropped slope, a
    mis-signed correction or a first-order corrector would not."""
    res = (32, 64, 128, 256)
    sols = [_evolve(_acoustic(N), N, 0.1, method, lim).reshape(5, N) for N in res]
    errs = []
    for a, b in zip(sols[:-1], sols[1:]):
        bc = 0.5 * (b[:, 0::2] + b[:, 1::2])
        errs.append(np.abs(a - bc).mean(axis=1))
    errs = np.array(errs)
    order = np.log2(errs[-2] / errs[-1])
    # energy and density and momentum all carry the wave; pz, py stay zero
    assert np.all(order[:3] > 1.7), (errs, order)
