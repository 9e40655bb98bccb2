"""Pins the CPU oracle's hydro arithmetic (oracle/imogen_oracle.c) against
closed forms, textbook formulations and invariants -- the reference ships no
golden vectors for the fluid step (its tests are run-time convergence tests,
run/test_suite/tsSod.m, and formula checks, gpuclass/unitTest.m).

What pins what:
  limiters          closed forms of gpuclass/limiter_functions.h
  HLLC flux+speeds  textbook HLLC/HLLE (tests/hydro_textbook.py), zero limiter
  MUSCL order       2nd-order L1 convergence on a smooth entropy wave
  full sweep        Sod shock tube vs the exact Riemann solution (~1st order)
  exact solver      Toro's published star states (Table 4.2)
  direction maps    X sweep of A == Y/Z sweep of the transposed A, bitwise
  sweep order       reference permutation semantics return to identity
  CFL, sound speed  numpy closed forms (gpuclass/unitTest.m:409-453)
  BCs               index maps of each kernel on labelled data
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


# ---------------------------------------------------------------- exact solver
@pytest.mark.parametrize("left,right,ps,us", [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.0),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975),
])
def test_exact_solver_toro_table(left, right, ps, us):
    p, u = er.star_state(left, right, 1.4)
    assert abs(p - ps) <= 6e-5 * max(1.0, abs(ps)) + 1e-5
    assert abs(u - us) <= 6e-5 * max(1.0, abs(us)) + 1e-5


# ---------------------------------------------------------------- limiters
def test_limiters_closed_form():
    rng = np.random.default_rng(1)
    for a, b in rng.normal(size=(200, 2)):
        ab = a * b
        exp = 0.75 * ab * (a + b) / (a * a + ab + b * b) if ab > 0 else 0.0
        assert po.slope_limiter(po.LIM_OSPRE, a, b) == pytest.approx(exp, rel=1e-14, abs=1e-300)
        exp = ab / (a + b) if ab > 0 else 0.0
        assert po.slope_limiter(po.LIM_VANLEER, a, b) == pytest.approx(exp, rel=1e-14, abs=1e-300)
        exp = 0.0 if ab < 0 else 0.5 * (a if abs(a) <= abs(b) else b)
        assert po.slope_limiter(po.LIM_MINMOD, a, b) == pytest.approx(exp, rel=1e-7, abs=1e-300)
        exp = 0.0 if ab <= 0 else a * min(0.25 + 0.25 * b / a, b / a, 1.0)
        assert po.slope_limiter(po.LIM_MC, a, b) == pytest.approx(exp, rel=1e-14, abs=1e-300)
        assert po.slope_limiter(po.LIM_ZERO, a, b) == 0.0
    # minmod goes through float (limiter_functions.h:70)
    assert po.slope_limiter(po.LIM_MINMOD, 0.1, 0.3) == float(np.float32(0.05))


# ---------------------------------------------------------------- HLLC textbook
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_first_order_sweep_matches_textbook_hllc(seed):
    """With the zero slope limiter one sweep is first-order HLLC + Heun; this
    pins the star states, the four-case selection and the Einfeldt speeds."""
    rng = np.random.default_rng(seed)
    n = 24
    rho = rng.uniform(0.2, 2.0, n)
    v = rng.uniform(-1.5, 1.5, (3, n))
    P = rng.uniform(0.1, 3.0, n)
    q = cons(rho, v[0], v[1], v[2], P).reshape(5, 1, 1, n).copy()
    lam = 0.1
    U = q.reshape(5, n).copy()
    po.fluid_update_1d(q, 1, lam, G, limiter=po.LIM_ZERO)
    ref = tb.first_order_heun_line(U, lam, G)
    err = np.abs(q.reshape(5, n) - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err


def test_supersonic_cases_match_textbook():
    """States moving faster than sound in both directions exercise the
    plain-flux cases (0 and 1)."""
    n = 16
    for vx in (4.0, -4.0):
        rng = np.random.default_rng(7)
        rho = rng.uniform(0.5, 1.5, n)
        P = rng.uniform(0.1, 0.3, n)
        q = cons(rho, np.full(n, vx), 0.1 * rng.normal(size=n), 0.1 * rng.normal(size=n), P)
        q = q.reshape(5, 1, 1, n).copy()
        U = q.reshape(5, n).copy()
        po.fluid_update_1d(q, 1, 0.05, G, limiter=po.LIM_ZERO)
        ref = tb.first_order_heun_line(U, 0.05, G)
        assert np.abs(q.reshape(5, n) - ref).max() / np.abs(ref).max() < 1e-13


# ---------------------------------------------------------------- invariants
def test_uniform_state_is_fixed_point():
    q = cons(np.full((6, 7, 9), 1.3), np.full((6, 7, 9), 0.4), np.full((6, 7, 9), -0.2),
             np.full((6, 7, 9), 0.1), np.full((6, 7, 9), 0.9))
    q0 = q.copy()
    for it in range(2):
        po.fluidstep(q, [0.1, 0.1, 0.1], 0.01, 1, it, G)
        po.fluidstep(q, [0.1, 0.1, 0.1], 0.01, -1, it, G)
    assert np.array_equal(q, q0)


def test_periodic_conservation():
    rng = np.random.default_rng(3)
    sh = (10, 12, 14)
    q = cons(rng.uniform(0.5, 2, sh), rng.normal(size=sh) * 0.3, rng.normal(size=sh) * 0.3,
             rng.normal(size=sh) * 0.3, rng.uniform(0.5, 2, sh))
    tot0 = q.reshape(5, -1).sum(axis=1)
    for it in range(3):
        po.fluidstep(q, [0.1, 0.1, 0.1], 0.004, 1, it, G)
        po.fluidstep(q, [0.1, 0.1, 0.1], 0.004, -1, it, G)
    tot1 = q.reshape(5, -1).sum(axis=1)
    scale = np.abs(q.reshape(5, -1)).sum(axis=1)
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * scale)


def _sweep(q, d, lam, lim=po.LIM_OSPRE):
    q = q.copy()
    po.fluid_update_1d(q, d, lam, G, limiter=lim)
    return q


def test_direction_maps_are_transposes():
    """An X sweep of A equals a Y (Z) sweep of A with x,y (x,z) exchanged --
    bitwise, since the reference reads the momenta in exactly that order."""
    rng = np.random.default_rng(5)
    sh = (7, 9, 11)  # (nz, ny, nx)
    q = cons(rng.uniform(0.5, 2, sh), rng.normal(size=sh), rng.normal(size=sh), rng.normal(size=sh),
             rng.uniform(0.5, 2, sh))
    rx = _sweep(q, 1, 0.05)
    # exchange x<->y: array axes (z,y,x)->(z,x,y), momentum slabs px<->py
    qy = np.ascontiguousarray(q.transpose(0, 1, 3, 2)[[0, 1, 3, 2, 4]])
    ry = _sweep(qy, 2, 0.05)
    assert np.array_equal(np.ascontiguousarray(ry.transpose(0, 1, 3, 2)[[0, 1, 3, 2, 4]]), rx)
    # exchange x<->z: axes (z,y,x)->(x,y,z), slabs px<->pz
    qz = np.ascontiguousarray(q.transpose(0, 3, 2, 1)[[0, 1, 4, 3, 2]])
    rz = _sweep(qz, 3, 0.05)
    assert np.array_equal(np.ascontiguousarray(rz.transpose(0, 3, 2, 1)[[0, 1, 4, 3, 2]]), rx)


def test_sweep_order_matches_reference_permutations():
    """Replays the reference's transpose codes (flux.cu:44-49; codes of
    cudaArrayRotateB.cu:551-640) on an axis tuple: each flux call must find
    its direction in memory-x and every sweep must end in XYZ order."""
    preperm = [0, 2, 0, 2, 3, 3]
    fluxcall = [[1, 2, 1, 2, 3, 3], [3, 1, 2, 3, 1, 2], [2, 3, 3, 1, 2, 1]]
    permcall = [[3, 2, 2, 3, 3, 2], [2, 3, 3, 5, 2, 6], [6, 3, 5, 0, 2, 0]]

    def apply(p, code):
        p = list(p)
        if code == 2:
            p[0], p[1] = p[1], p[0]
        elif code == 3:
            p[0], p[2] = p[2], p[0]
        elif code == 4:
            p[1], p[2] = p[2], p[1]
        elif code == 5:
            p = [p[1], p[2], p[0]]
        elif code == 6:
            p = [p[2], p[0], p[1]]
        return p

    for sweep in range(6):
        p = apply([1, 2, 3], preperm[sweep])
        order = []
        for n in range(3):
            assert p[0] == fluxcall[n][sweep]
            order.append(p[0])
            p = apply(p, permcall[n][sweep])
        assert p == [1, 2, 3]
        assert tuple(order) == po.sweep_order(sweep)
    # forward call of iteration k uses sweep (k+3)%6, backward k%6 (flux.cu:32)
    assert po.sweep_order((0 + 3) % 6) == (2, 3, 1)
    assert po.sweep_order(0) == (1, 3, 2)


# ---------------------------------------------------------------- convergence
def _run_1d(q, h, tend, method=po.HLLC, lim=po.LIM_OSPRE, cfl=0.4, modes=(po.BC_CONST,) * 6):
    t, it = 0.0, 0
    N = q.shape[-1]
    while t < tend * (1 - 1e-14):
        cs = po.soundspeed(q, G)
        tau = po.cfl_timestep(q, cs, [h, 1, 1], method, (N, 1, 1))
        dt = cfl * tau
        if t + 2 * dt > tend:  # TimeManager.m:180-183
            dt = 0.5 * (tend - t)
        po.fluidstep(q, [h, 1, 1], dt, 1, it, G, method, lim, modes=modes)
        po.fluidstep(q, [h, 1, 1], dt, -1, it, G, method, lim, modes=modes)
        t += 2 * dt
        it += 1
    return q


def test_sod_converges_to_exact_solution():
    errs = []
    for N in (100, 200, 400):
        dx = 1.0 / N
        x = (np.arange(N) + 0.5) * dx
        q = cons(np.where(x < 0.5, 1.0, 0.125), 0 * x, 0 * x, 0 * x, np.where(x < 0.5, 1.0, 0.1))
        q = q.reshape(5, 1, 1, N).copy()
        _run_1d(q, dx, 0.2)
        ex = er.cell_averages(x, dx, 0.5, 0.2, (1, 0, 1), (0.125, 0, 0.1), G)
        errs.append(np.abs(q[0, 0, 0] - ex[0]).mean())
    assert errs[0] < 6e-3 and errs[2] < 2e-3, errs
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 0.75), orders


def test_smooth_wave_second_order():
    """Entropy wave rho = 1 + 0.2 sin(2 pi (x - t)), u = 1, P = 1 on a periodic
    line: a MUSCL scheme converges at ~2nd order in L1 (a first-order or
    broken reconstruction would not)."""
    errs = []
    for N in (32, 64, 128):
        dx = 1.0 / N
        x = (np.arange(N) + 0.5) * dx
        # exact cell averages of rho at t = 0 and t = tend
        def avg(t):
            return 1 + 0.2 * (np.cos(2 * np.pi * (x - dx / 2 - t)) - np.cos(2 * np.pi * (x + dx / 2 - t))) / (2 * np.pi * dx)
        q = cons(avg(0.0), np.ones(N), 0 * x, 0 * x, np.ones(N)).reshape(5, 1, 1, N).copy()
        _run_1d(q, dx, 0.25, modes=(po.BC_CIRCULAR,) * 6)
        errs.append(np.abs(q[0, 0, 0] - avg(0.25)).mean())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert orders[-1] > 1.6, (errs, orders)


# ---------------------------------------------------------------- CFL / sound speed
def test_soundspeed_and_cfl_closed_form():
    rng = np.random.default_rng(11)
    sh = (5, 6, 7)
    q = cons(rng.uniform(0.5, 2, sh), rng.normal(size=sh), rng.normal(size=sh), rng.normal(size=sh),
             rng.uniform(0.5, 2, sh))
    cs = po.soundspeed(q, G)
    ke = 0.5 * (q[2] ** 2 + q[3] ** 2 + q[4] ** 2) / q[0]
    exp_cs = np.sqrt(np.maximum(G * (G - 1) * (q[1] - ke) / q[0], 0))
    assert np.allclose(cs, exp_cs, rtol=1e-14, atol=0)
    h = [0.1, 0.07, 0.13]
    vmag = np.sqrt(q[2] ** 2 + q[3] ** 2 + q[4] ** 2) / q[0]
    tau = po.cfl_timestep(q, cs, h, po.HLLC)
    assert tau == pytest.approx(0.5 * np.min(min(h) / (vmag + cs)), rel=1e-14)
    tau_xj = po.cfl_timestep(q, cs, h, po.XINJIN)
    exp = np.min(np.minimum(np.minimum(h[0] / (np.abs(q[2]) / q[0] + cs), h[1] / (np.abs(q[3]) / q[0] + cs)),
                            h[2] / (np.abs(q[4]) / q[0] + cs)))
    assert tau_xj == pytest.approx(exp, rel=1e-14)
    # 2D: only x,y momenta and spacings count (cflTimestep.cu:266-275)
    q2 = q[:, :1].copy()
    cs2 = po.soundspeed(q2, G)
    v2 = np.sqrt(q2[2] ** 2 + q2[3] ** 2) / q2[0]
    assert po.cfl_timestep(q2, cs2, h, po.HLL) == pytest.approx(0.5 * np.min(min(h[:2]) / (v2 + cs2)), rel=1e-14)


# ---------------------------------------------------------------- boundary conditions
def _labelled(sh):
    q = np.zeros((5,) + sh)
    for s in range(5):
        q[s] = np.arange(np.prod(sh)).reshape(sh) + 1000 * s + 1
    return q


def test_bc_const_and_mirror_index_maps():
    d = 4
    # x direction on (nz, ny, nx) = (2, 3, 16)
    q = _labelled((2, 3, 16))
    r = q.copy()
    po.set_fluid_boundary(r, 1, [po.BC_CONST] * 6, d)
    assert np.all(r[..., :d] == q[..., d:d + 1])
    assert np.all(r[..., -d:] == q[..., -d - 1:-d])  # last interior cell
    r = q.copy()
    po.set_fluid_boundary(r, 1, [po.BC_MIRROR] * 6, d)
    for g in range(d):
        sign = np.array([1, 1, -1, 1, 1]).reshape(5, 1, 1)  # px flips on x faces
        assert np.all(r[..., d - 1 - g] == sign * q[..., d + 1 + g])
        assert np.all(r[..., 16 - d + g] == sign * q[..., 16 - d - 2 - g])
    # y direction: +y const copies the innermost ghost (cudaStatics.cu:1461-1466)
    q = _labelled((2, 16, 3))
    r = q.copy()
    po.set_fluid_boundary(r, 2, [po.BC_CONST] * 6, d)
    assert np.all(r[:, :, :d, :] == q[:, :, d:d + 1, :])
    assert np.all(r[:, :, -d:, :] == q[:, :, -d:-d + 1, :])
    # z direction: z-minus mirror reflects about 4.5 (cudaStatics.cu:1496-1500)
    q = _labelled((16, 2, 3))
    r = q.copy()
    po.set_fluid_boundary(r, 3, [po.BC_MIRROR] * 6, d)
    for g in range(d):
        assert np.all(r[[0, 1, 2, 3], g] == q[[0, 1, 2, 3], 2 * d + 1 - g])
        assert np.all(r[4, g] == -q[4, 2 * d + 1 - g])
        assert np.all(r[[0, 1, 2, 3], 16 - 1 - g] == q[[0, 1, 2, 3], 16 - 1 - 2 * d + g])


@pytest.mark.parametrize("direction", [2, 3])
@pytest.mark.parametrize("mode", ["const", "mirror", "linear"])
def test_swept_boundaries_are_x_boundaries_of_the_rotated_arrays(direction, mode):
    """flux.cu rotates the arrays so the swept direction is memory x and then
    calls setFluidBoundary (flux.cu:99-107), which picks the memory-x kernels:
    the fluid step's boundaries along y or z are the x boundaries of the
    transposed arrays, with that direction's momentum as the normal one."""
    d = 4
    sh = (13, 14, 15) if direction == 3 else (3, 14, 5)
    q = _labelled(sh)
    code = {"const": po.BC_CONST, "mirror": po.BC_MIRROR, "linear": po.BC_LINEAR}[mode]
    modes = [po.BC_CIRCULAR] * 6
    modes[2 * (direction - 1)] = code
    if mode != "linear":  # linear extrapolation exists on the -x face only
        modes[2 * (direction - 1) + 1] = code
    r = q.copy()
    po.fluid_boundary_swept(r, direction, modes, d)
    # (slab, z, y, x): move the swept axis last (memory x), and the swept
    # direction's momentum into the px slot
    ax = 1 if direction == 3 else 2
    order = [0, 1, 2, 3]
    order[ax], order[3] = order[3], order[ax]
    t = np.ascontiguousarray(np.transpose(q, order))
    slabs = [0, 1, 1 + direction] + [k for k in (2, 3, 4) if k != 1 + direction]
    t = np.ascontiguousarray(t[slabs])
    xm = [code, po.BC_CIRCULAR if mode == "linear" else code] + [po.BC_CIRCULAR] * 4
    po.set_fluid_boundary(t, 1, xm, d)
    back = np.empty_like(t)
    back[slabs] = t
    expect = np.transpose(back, order)
    assert np.array_equal(r, expect)
    # and the per-axis kernels differ where the quirks act (const: +face
    # source; mirror along z: the -face offset), so the choice matters
    if mode == "const" or (mode == "mirror" and direction == 3):
        r2 = q.copy()
        po.set_fluid_boundary(r2, direction, modes, d)
        assert not np.array_equal(r2, r)
