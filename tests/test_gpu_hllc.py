"""GPU parity of the HLLC fluid step (libimogen_hip.so) against the CPU oracle.

Bar: relative L1 <= 1e-12 on every conserved variable per step, the tolerance
the fluid-step contract states for FP64; per-cell error is also bounded
(1e-10 of the variable's scale) so that a localised indexing bug cannot hide in
the L1 average. The GPU kernels use FMA contraction, the oracle does not, so
results agree to rounding, not bitwise.
"""
import numpy as np
import pytest

from oracle import exact_riemann as er
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

G = 5.0 / 3.0
L1_TOL = 1e-12
CELL_TOL = 1e-10


def gc():
    from gpuimogen_amd import gpuclass
    return gpuclass


def abi():
    from gpuimogen_amd import abi as a
    return a


def cons(rho, vx, vy, vz, P, g=G):
    q = np.zeros((5,) + np.shape(rho))
    q[0] = rho
    q[1] = P / (g - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    q[2] = rho * vx
    q[3] = rho * vy
    q[4] = rho * vz
    return q


def random_state(sh, seed, smooth=False):
    rng = np.random.default_rng(seed)
    if smooth:
        nz, ny, nx = sh
        z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
        ph = rng.uniform(0, 2 * np.pi, 6)
        rho = 1 + 0.3 * np.sin(2 * np.pi * x + ph[0]) * np.cos(2 * np.pi * y + ph[1]) + 0.2 * np.sin(2 * np.pi * z + ph[2])
        P = 1 + 0.3 * np.cos(2 * np.pi * (x + y + z) + ph[3])
        vx = 0.5 * np.sin(2 * np.pi * y + ph[4])
        vy = 0.5 * np.cos(2 * np.pi * z + ph[5])
        vz = 0.3 * np.sin(2 * np.pi * x)
        return cons(rho, vx, vy, vz, P)
    rho = rng.uniform(0.3, 2.0, sh)
    P = rng.uniform(0.2, 2.0, sh)
    v = rng.uniform(-1.0, 1.0, (3,) + sh)
    return cons(rho, v[0], v[1], v[2], P)


def compare(got, ref, what):
    # The van Albada limiter of the reference evaluates 0/0 for two zero
    # differences (limiter_functions.h:101-106); NaNs must then appear in the
    # same cells on both sides, and the finite cells must agree.
    nan_g, nan_r = np.isnan(got), np.isnan(ref)
    assert np.array_equal(nan_g, nan_r), (what, "NaN pattern differs")
    got = np.where(nan_r, 0.0, got)
    ref = np.where(nan_r, 0.0, ref)
    for s in range(5):
        d = np.abs(got[s] - ref[s])
        scale = np.abs(ref[s]).sum()
        rel_l1 = d.sum() / scale if scale > 0 else d.sum()
        cell = d.max() / max(np.abs(ref[s]).max(), 1e-300)
        print(f"{what} slab {s}: rel L1 {rel_l1:.3e}  max cell {cell:.3e}", flush=True)
        assert rel_l1 <= L1_TOL, (what, s, rel_l1)
        assert cell <= CELL_TOL, (what, s, cell)


def geom(sh, h=(0.01, 0.012, 0.014)):
    nz, ny, nx = sh
    return gc().GeometryManager((nx, ny, nz), h)


def test_upload_download_and_slabs():
    g = gc()
    rng = np.random.default_rng(0)
    a = rng.normal(size=(7, 5, 9))
    t = g.GPU_Type(a)
    assert np.array_equal(t.array, a)
    t.createSlabs(3)
    assert np.array_equal(t.slab(0).array, a)
    b = rng.normal(size=a.shape)
    t.setslab(2, b)
    assert np.array_equal(t.slab(2).array, b)
    c = g.GPU_Type(dims=(9, 5, 7))
    assert c.dims == (9, 5, 7)
    t.free()
    c.free()


def test_gpu_ctrl_reset_and_info():
    """GPU_ctrl('reset') releases the cached temporaries between calls without
    touching the caller's arrays: a fluid step after a reset gives bitwise the
    step without one. GPU_ctrl('info') names the device."""
    import ctypes
    a = abi()
    sh = (10, 12, 14)
    q = random_state(sh, seed=21, smooth=True)
    gm = geom(sh, (0.05, 0.05, 0.05))
    fa = gc().FluidManager(q, G, bc_modes=(a.extrapConstant,) * 6)
    fb = gc().FluidManager(q, G, bc_modes=(a.extrapConstant,) * 6)
    for it in range(2):
        for direction in (1, -1):
            gc().fluidstep([fa], None, None, None, [0.002, 1, direction, it, 2], gm)
            a.call("GPU_ctrl_reset")
            gc().fluidstep([fb], None, None, None, [0.002, 1, direction, it, 2], gm)
    assert np.array_equal(fa.state, fb.state)
    fa.free()
    fb.free()
    buf = ctypes.create_string_buffer(512)
    a.call("GPU_ctrl_info", 0, buf, 512)
    text = buf.value.decode()
    assert text.startswith("device 0:") and "compute units" in text
    print(text, flush=True)


# the last six: X lines of 300-512 cells, swept as two halves by a wave pair
# (3, 4 and 5 cells per lane, both edges of the 3- and 4-cell classes, odd
# lengths) or, at exactly 384 and 512 cells, whole with 6 and 8 cells per lane
SHAPES = [(23, 29, 37), (4, 6, 256), (6, 5, 192), (5, 200, 8), (150, 3, 20), (1, 40, 48), (1, 1, 100),
          (3, 7, 97), (3, 4, 300), (2, 3, 368), (2, 3, 384), (2, 3, 401), (2, 2, 497), (2, 2, 512)]


@pytest.mark.parametrize("sh", SHAPES)
@pytest.mark.parametrize("direction", [1, 2, 3])
def test_single_sweep_parity(sh, direction):
    nz, ny, nx = sh
    n = (nx, ny, nz)[direction - 1]
    if n < 4:
        pytest.skip("direction not swept (fewer than 4 cells)")
    q = random_state(sh, seed=hash((sh, direction)) % 1000)
    print(f"case shape={sh} dir={direction}", flush=True)
    f = gc().FluidManager(q, G, bc_modes=(abi().circular,) * 6)
    gm = geom(sh)
    dt = 0.2 * gm.h[direction - 1]
    gc().cudaFluidStep([dt, 1, G, 1e-8, 2, direction], f, gm)
    got = f.state
    ref = q.copy()
    po.fluid_update_1d(ref, direction, dt / gm.h[direction - 1], G)
    compare(got, ref, f"sweep{direction} {sh}")
    f.free()


def limiter_case(limiter, seed):
    """State for a limiter parity test and whether each direction must start
    afresh from it. The continuous limiters run on smooth data, three chained
    sweeps. Two reference limiters are discontinuous, so their results hang on
    the last bit of the primitive differences: superbee returns a full (not
    half) slope on one branch (limiter_functions.h:80-86), which on smooth data
    amplifies rounding noise ~1e11-fold over chained sweeps in the oracle
    itself, and van Albada evaluates 0/0 = NaN for two zero differences
    (:101-106), which on smooth data appears wherever rounding happens to give
    exact zeros. Both run on rough data, one sweep per direction; van Albada
    also gets a block of identical cells, where the differences are exactly
    zero whatever the rounding, so its NaN case is exercised."""
    sh = (12, 14, 16)
    if limiter not in (3, 4):
        return random_state(sh, seed=seed, smooth=True), False
    q = random_state(sh, seed=seed)
    if limiter == 4:
        q[:, 3:9, 4:10, 5:11] = q[:, 3:4, 4:5, 5:6]
    return q, True


@pytest.mark.parametrize("limiter", range(7))
def test_limiters_parity(limiter):
    q, fresh = limiter_case(limiter, seed=40 + limiter)
    sh = q.shape[1:]
    f = gc().FluidManager(q, G)
    gm = geom(sh, (0.1, 0.1, 0.1))
    ref = q.copy()
    for d in (1, 2, 3):
        if fresh:
            f.set_state(q)
            ref = q.copy()
        gc().cudaFluidStep([0.02, 1, G, 1e-8, 2, d], f, gm, limiter=limiter)
        po.fluid_update_1d(ref, d, 0.2, G, limiter=limiter)
        compare(f.state, ref, f"limiter {limiter} dir {d}")
    f.free()


@pytest.mark.parametrize("bc", ["const", "mirror", "circular"])
@pytest.mark.parametrize("sh", [(20, 24, 28), (1, 36, 44), (1, 1, 64), (12, 11, 13), (3, 12, 128), (2, 3, 256),
                                (3, 12, 32), (3, 12, 320), (2, 12, 451), (10, 12, 256)])
def test_fluidstep_parity_with_boundaries(bc, sh):
    """(12, 11, 13): lines of 11-13 cells, where a mirror face reads cells the
    opposite face writes, so the faces must be set one after the other.
    (3, 12, 128), (2, 3, 256), (3, 12, 32): whole X lines of 2, 4 and 1 cells
    per lane (the last with two lines per wave) between swept y lines and
    unswept z lines. (10, 12, 256): 4-cell lanes over whole X lines with y
    and z swept too."""
    a = abi()
    mode = {"const": a.extrapConstant, "mirror": a.mirror, "circular": a.circular}[bc]
    omode = {"const": po.BC_CONST, "mirror": po.BC_MIRROR, "circular": po.BC_CIRCULAR}[bc]
    q = random_state(sh, seed=7, smooth=True)
    f = gc().FluidManager(q, G, bc_modes=(mode,) * 6)
    gm = geom(sh, (0.05, 0.05, 0.05))
    ref = q.copy()
    for it in range(3):
        tau = gc().cflTimestepFused(f, gm, 2)
        cs = po.soundspeed(ref, G)
        tau_ref = po.cfl_timestep(ref, cs, gm.h, po.HLLC, gm.globalDomainRez)
        assert abs(tau - tau_ref) <= 1e-13 * tau_ref
        dt = 0.4 * tau_ref
        for direction in (1, -1):
            gc().fluidstep([f], None, None, None, [dt, 1, direction, it, 2], gm)
            po.fluidstep(ref, gm.h, dt, direction, it, G, po.HLLC, po.LIM_OSPRE, modes=(omode,) * 6)
    compare(f.state, ref, f"fluidstep {bc} {sh}")
    f.free()


@pytest.mark.parametrize("bc", ["const", "mirror"])
def test_set_fluid_boundary_on_short_lines(bc):
    """setFluidBoundary on unrotated arrays (each memory axis's own kernels,
    with the +y/+z const and z-minus mirror quirks) on 11-13-cell lines, whose
    two faces overlap for mirror: exactly the oracle's face-by-face result."""
    a = abi()
    sh = (12, 11, 13)
    q = random_state(sh, seed=11, smooth=True)
    mode = a.extrapConstant if bc == "const" else a.mirror
    omode = po.BC_CONST if bc == "const" else po.BC_MIRROR
    f = gc().FluidManager(q, G, bc_modes=(mode,) * 6)
    gm = geom(sh, (0.05, 0.05, 0.05))
    ref = q.copy()
    for d in (1, 2, 3):
        gc().setBoundaries(f, gm, d)
        po.set_fluid_boundary(ref, d, (omode,) * 6)
        assert np.array_equal(f.state, ref), d
    f.free()


def test_cfl_and_soundspeed_parity():
    sh = (10, 11, 12)
    q = random_state(sh, seed=3)
    f = gc().FluidManager(q, G)
    gm = geom(sh)
    cs = gc().cudaSoundspeed(f, G)
    cs_ref = po.soundspeed(q, G)
    assert np.allclose(cs.array, cs_ref, rtol=1e-14, atol=0)
    for method in (1, 2, 3):
        tau = gc().cflTimestep(f, cs, gm, method)
        tau_ref = po.cfl_timestep(q, cs_ref, gm.h, method, gm.globalDomainRez)
        assert tau == pytest.approx(tau_ref, rel=1e-14)
        assert gc().cflTimestepFused(f, gm, method) == pytest.approx(tau_ref, rel=1e-14)
    cs.free()
    f.free()


def test_sod_on_gpu_matches_exact_solution():
    """tsSod-style check (run/test_suite/tsSod.m): HLLC Sod at 400 cells to
    t = 0.2 against the exact Riemann solution, and against the oracle run."""
    N = 400
    dx = 1.0 / N
    x = (np.arange(N) + 0.5) * dx
    g = 1.4
    q = cons(np.where(x < 0.5, 1.0, 0.125), 0 * x, 0 * x, 0 * x, np.where(x < 0.5, 1.0, 0.1), g)
    q = q.reshape(5, 1, 1, N)
    a = abi()
    f = gc().FluidManager(q, g, bc_modes=(a.extrapConstant,) * 6)
    gm = gc().GeometryManager((N, 1, 1), (dx, 1, 1))
    ref = q.copy()
    t, it = 0.0, 0
    while t < 0.2 * (1 - 1e-14):
        tau = gc().cflTimestepFused(f, gm, 2)
        dt = 0.4 * tau
        if t + 2 * dt > 0.2:
            dt = 0.5 * (0.2 - t)
        for direction in (1, -1):
            gc().fluidstep([f], None, None, None, [dt, 1, direction, it, 2], gm)
            po.fluidstep(ref, gm.h, dt, direction, it, g, po.HLLC, modes=(po.BC_CONST,) * 6)
        t += 2 * dt
        it += 1
    got = f.state
    ex = er.cell_averages(x, dx, 0.5, 0.2, (1, 0, 1), (0.125, 0, 0.1), g)
    err = np.abs(got[0, 0, 0] - ex[0]).mean()
    print(f"Sod N={N}: L1(rho) vs exact = {err:.3e} after {it} iterations", flush=True)
    assert err < 2e-3
    # whole-run agreement with the oracle (errors of both stay at rounding level)
    scale = np.maximum(np.abs(ref).reshape(5, -1).sum(axis=1), np.abs(ref).sum() * 1e-3)
    d = np.abs(got - ref).reshape(5, -1).sum(axis=1) / scale
    print("Sod whole-run rel L1 vs oracle:", d, flush=True)
    assert np.all(d < 1e-10)
    f.free()
