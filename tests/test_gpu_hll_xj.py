"""GPU parity of the HLL and Xin-Jin fluid steps (libimogen_hip.so) against
the CPU oracle, to the same bar as the HLLC path: relative L1 <= 1e-12 per
conserved variable per step and 1e-10 of the variable's scale per cell."""
import numpy as np
import pytest

from oracle import exact_riemann as er
from oracle import pyoracle as po

from .test_gpu_hllc import G, SHAPES, abi, compare, cons, gc, geom, limiter_case, random_state

pytestmark = pytest.mark.gpu

METHODS = {"hll": (1, po.HLL), "xinjin": (3, po.XINJIN)}


@pytest.mark.parametrize("method", ["hll", "xinjin"])
@pytest.mark.parametrize("sh", SHAPES)
@pytest.mark.parametrize("direction", [1, 2, 3])
def test_single_sweep_parity(method, sh, direction):
    m, om = METHODS[method]
    nz, ny, nx = sh
    n = (nx, ny, nz)[direction - 1]
    if n < 4:
        pytest.skip("direction not swept (fewer than 4 cells)")
    q = random_state(sh, seed=hash((sh, direction, method)) % 1000)
    print(f"case {method} shape={sh} dir={direction}", flush=True)
    f = gc().FluidManager(q, G)
    gm = geom(sh)
    dt = 0.15 * gm.h[direction - 1]
    gc().cudaFluidStep([dt, 1, G, 1e-8, m, direction], f, gm)
    ref = q.copy()
    po.fluid_update_1d(ref, direction, dt / gm.h[direction - 1], G, method=om)
    compare(f.state, ref, f"{method} sweep{direction} {sh}")
    f.free()


@pytest.mark.parametrize("limiter", range(7))
def test_hll_limiters_parity(limiter):
    q, fresh = limiter_case(limiter, seed=60 + limiter)
    sh = q.shape[1:]
    f = gc().FluidManager(q, G)
    gm = geom(sh, (0.1, 0.1, 0.1))
    ref = q.copy()
    for d in (1, 2, 3):
        if fresh:
            f.set_state(q)
            ref = q.copy()
        gc().cudaFluidStep([0.02, 1, G, 1e-8, 1, d], f, gm, limiter=limiter)
        po.fluid_update_1d(ref, d, 0.2, G, method=po.HLL, limiter=limiter)
        compare(f.state, ref, f"hll limiter {limiter} dir {d}")
    f.free()


@pytest.mark.parametrize("method", ["hll", "xinjin"])
def test_floors_parity(method):
    """A block of cells below the density floor and a block with internal
    energy below rho*emin exercise the floors of both methods in every
    direction (each direction starts from the same state)."""
    m, om = METHODS[method]
    sh = (9, 10, 70)
    rng = np.random.default_rng(21)
    rho = rng.uniform(0.5, 2.0, sh)
    rho[:, :, 40:52] = 0.2
    P = rng.uniform(0.2, 2.0, sh)
    P[:, :, 15:30] = 1e-7
    v = rng.uniform(-0.5, 0.5, (3,) + sh)
    v[:, :, :, 15:30] *= 1e-3
    q = cons(rho, v[0], v[1], v[2], P)
    rhomin = 0.3
    f = gc().FluidManager(q, G, minmass=rhomin)
    gm = geom(sh, (0.1, 0.1, 0.1))
    for d in (1, 2, 3):
        f.set_state(q)
        ref = q.copy()
        gc().cudaFluidStep([0.005, 1, G, rhomin, m, d], f, gm)
        po.fluid_update_1d(ref, d, 0.05, G, method=om, rhomin=rhomin)
        compare(f.state, ref, f"{method} floors dir {d}")
    f.free()


def test_hll_isothermal_parity():
    sh = (6, 20, 36)
    ciso = 0.7
    q = random_state(sh, seed=5)
    f = gc().FluidManager(q, 2.0, cisothermal=ciso)
    gm = geom(sh, (0.1, 0.1, 0.1))
    ref = q.copy()
    for d in (1, 2, 3):
        gc().cudaFluidStep([0.01, 1, 2.0, 1e-8, 1, d, ciso], f, gm)
        po.fluid_update_1d(ref, d, 0.1, 2.0, method=po.HLL, ciso=ciso)
        compare(f.state, ref, f"hll isothermal dir {d}")
    f.free()


@pytest.mark.parametrize("method", ["hll", "xinjin"])
@pytest.mark.parametrize("bc", ["const", "mirror"])
def test_fluidstep_parity_with_boundaries(method, bc):
    m, om = METHODS[method]
    a = abi()
    mode = {"const": a.extrapConstant, "mirror": a.mirror}[bc]
    omode = {"const": po.BC_CONST, "mirror": po.BC_MIRROR}[bc]
    sh = (20, 24, 28)
    q = random_state(sh, seed=17, smooth=True)
    f = gc().FluidManager(q, G, bc_modes=(mode,) * 6)
    gm = geom(sh, (0.05, 0.05, 0.05))
    ref = q.copy()
    for it in range(3):
        tau = gc().cflTimestepFused(f, gm, m)
        cs = po.soundspeed(ref, G)
        tau_ref = po.cfl_timestep(ref, cs, gm.h, om, gm.globalDomainRez)
        assert abs(tau - tau_ref) <= 1e-13 * tau_ref
        dt = 0.4 * tau_ref
        for direction in (1, -1):
            gc().fluidstep([f], None, None, None, [dt, 1, direction, it, m], gm)
            po.fluidstep(ref, gm.h, dt, direction, it, G, om, po.LIM_OSPRE, modes=(omode,) * 6)
    compare(f.state, ref, f"{method} fluidstep {bc}")
    f.free()


@pytest.mark.parametrize("method", ["hll", "xinjin"])
def test_sod_on_gpu(method):
    m, om = METHODS[method]
    N = 400
    dx = 1.0 / N
    x = (np.arange(N) + 0.5) * dx
    g = 1.4
    q = cons(np.where(x < 0.5, 1.0, 0.125), 0 * x, 0 * x, 0 * x, np.where(x < 0.5, 1.0, 0.1), g)
    q = q.reshape(5, 1, 1, N)
    f = gc().FluidManager(q, g, bc_modes=(abi().extrapConstant,) * 6)
    gm = gc().GeometryManager((N, 1, 1), (dx, 1, 1))
    ref = q.copy()
    t, it = 0.0, 0
    while t < 0.2 * (1 - 1e-14):
        tau = gc().cflTimestepFused(f, gm, m)
        dt = 0.4 * tau
        if t + 2 * dt > 0.2:
            dt = 0.5 * (0.2 - t)
        for direction in (1, -1):
            gc().fluidstep([f], None, None, None, [dt, 1, direction, it, m], gm)
            po.fluidstep(ref, gm.h, dt, direction, it, g, om, modes=(po.BC_CONST,) * 6)
        t += 2 * dt
        it += 1
    got = f.state
    ex = er.cell_averages(x, dx, 0.5, 0.2, (1, 0, 1), (0.125, 0, 0.1), g)
    err = np.abs(got[0, 0, 0] - ex[0]).mean()
    print(f"{method} Sod N={N}: L1(rho) vs exact = {err:.3e} after {it} iterations", flush=True)
    assert err < 3e-3
    scale = np.maximum(np.abs(ref).reshape(5, -1).sum(axis=1), np.abs(ref).sum() * 1e-3)
    d = np.abs(got - ref).reshape(5, -1).sum(axis=1) / scale
    print(f"{method} Sod whole-run rel L1 vs oracle:", d, flush=True)
    assert np.all(d < 1e-10)
    f.free()
