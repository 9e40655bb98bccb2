"""The ensemble's CFL minimum folded into the last sweep of an iteration.

An iteration of the ensemble takes its time step from the minima that the
previous iteration's last sweep left for problems without dust whose sides
exceed 8 cells (the first iteration of each run call reduces the whole state).
That folded minimum must equal the full reduction bitwise, so a run of several
iterations in one call must match the same iterations run one call at a time
bit for bit, and both must match each problem advanced alone by the
single-array path. Rectangular problems put the folded range on both axes:
sides of 8 cells (no fold), 9 (one kept cell beyond the ghosts) and lines of
every X class."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (nx, ny)
SHAPES = [(8, 8), (9, 9), (9, 40), (40, 9), (8, 100), (100, 8), (64, 64), (300, 17), (17, 300), (509, 12),
          (12, 509), (130, 77), (385, 10)]


def rect_state(nx, ny, seed, gamma=1.4):
    """A jump in density and pressure across a diagonal plus smooth velocity
    fields, so that the CFL minimum sits at a different cell in each problem."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    k = rng.uniform(1, 4, size=4)
    ph = rng.uniform(0, 2 * np.pi, size=4)
    hi = (x + rng.uniform(0.5, 1.5) * y) > rng.uniform(0.4, 1.2)
    rho = np.where(hi, 2.0, 0.6) + 0.1 * np.sin(2 * np.pi * k[0] * x + ph[0])
    P = np.where(hi, 3.0, 0.4) * (1 + 0.05 * np.cos(2 * np.pi * k[1] * y + ph[1]))
    vx = 0.4 * np.sin(2 * np.pi * k[2] * y + ph[2])
    vy = -0.3 * np.cos(2 * np.pi * k[3] * x + ph[3])
    q = np.empty((5, ny, nx))
    q[0] = rho
    q[2] = rho * vx
    q[3] = rho * vy
    q[4] = 0.0
    q[1] = P / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy)
    return q


def single_path(state, iterations, nx, ny, limiter=0):
    from gpuimogen_amd import abi, gpuclass as gc
    f = gc.FluidManager(state.reshape(5, 1, ny, nx), 1.4, bc_modes=(abi.extrapConstant,) * 6)
    gm = gc.GeometryManager((nx, ny, 1), (1.0 / nx, 1.0 / ny, 1.0))
    t = 0.0
    for it in range(iterations):
        dt = 0.4 * gc.cflTimestepFused(f, gm, abi.METHOD_HLLC)
        gc.fluidstep([f], None, None, None, [dt, 1, 1, it, abi.METHOD_HLLC], gm, limiter=limiter)
        gc.fluidstep([f], None, None, None, [dt, 1, -1, it, abi.METHOD_HLLC], gm, limiter=limiter)
        t += 2 * dt
    out = f.state.reshape(5, ny, nx)
    f.free()
    return out, t


def build(states, limiter=0):
    from gpuimogen_amd import ensemble as ens
    nx = np.array([s[0] for s in SHAPES])
    ny = np.array([s[1] for s in SHAPES])
    e = ens.Ensemble(nx, ny, 1.0 / nx, 1.0 / ny, limiter=limiter)
    for p, s in enumerate(states):
        e.set_state(p, s)
    return e


@pytest.mark.parametrize("graph", [False, True])
def test_folded_minimum_equals_full_reduction(graph):
    states = [rect_state(nx, ny, 11 + p) for p, (nx, ny) in enumerate(SHAPES)]
    iters = 7  # every sweep order, each ending on X or Y
    a, b = build(states), build(states)
    print(f"one call of {iters} iterations vs {iters} calls, graph={graph}", flush=True)
    a.run(iters, graph=graph)
    for _ in range(iters):
        b.run(1, graph=graph)
    assert np.array_equal(a.times(), b.times())
    for p in range(len(SHAPES)):
        assert np.array_equal(a.state(p), b.state(p)), SHAPES[p]
    a.free()
    b.free()


# Ospre and van Leer have kernels of their own, minmod takes the generic one
@pytest.mark.parametrize("lim", [0, 1, 2])
def test_rectangular_problems_match_single_path(lim):
    states = [rect_state(nx, ny, 40 + p) for p, (nx, ny) in enumerate(SHAPES)]
    iters = 4
    e = build(states, limiter=lim)
    e.run(iters, graph=True)
    times = e.times()
    for p, (nx, ny) in enumerate(SHAPES):
        print(f"problem {p}: {nx} x {ny}, limiter {lim}", flush=True)
        ref, tref = single_path(states[p], iters, nx, ny, lim)
        got = e.state(p)
        for s in range(5):
            d = np.abs(got[s] - ref[s]).sum()
            scale = max(np.abs(ref[s]).sum(), 1e-300)
            assert d / scale <= 1e-12, (p, s, d / scale)
        assert times[p] == pytest.approx(tref, rel=1e-13)
    e.free()

