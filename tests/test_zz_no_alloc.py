"""performFluidUpdate_3D_noalloc (the allocation-free fluid step) against
performFluidUpdate_3D, called through the C ABI on the same inputs.

Two calls in a row, the second on a smaller grid, so that the second runs on
what the first left in the fixed scratch array; each must equal the allocating
entry point's result bitwise (the same kernels run on the same inputs, only
their scratch differs). Also: a grid whose sweeps need more than the fixed
array is rejected before any launch, with the state untouched.
"""
import ctypes

import numpy as np
import pytest

from gpuimogen_amd import abi, gpuclass as gc

pytestmark = pytest.mark.gpu

G = 5.0 / 3.0


def state(shape, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.3, 2.0, shape)
    P = rng.uniform(0.2, 2.0, shape)
    v = rng.uniform(-1.0, 1.0, (3,) + shape)
    q = np.zeros((5,) + shape)
    q[0] = rho
    q[1] = P / (G - 1) + 0.5 * rho * (v ** 2).sum(axis=0)
    q[2:5] = rho * v
    return q


def topo():
    t = abi.ParallelTopology()
    for i in range(3):
        t.nproc[i] = 1
    return t


def step(q, method, entry, modes, step_dir, step_num):
    nz, ny, nx = q.shape[1:]
    f = gc.FluidManager(q, G, bc_modes=modes)
    gm = gc.GeometryManager((nx, ny, nz), (0.011, 0.013, 0.017))
    p = gc.step_params(0.002, step_dir, step_num, method, gm, G, f.MINMASS, -1.0, abi.LIMITER_VANLEER)
    t = topo()
    arr = f.arrays()
    if entry == "noalloc":
        abi.call("performFluidUpdate_3D_noalloc", arr, ctypes.byref(t), p)
    else:
        abi.call("performFluidUpdate_3D", arr, ctypes.byref(t), p, None)
    out = f.state
    f.free()
    return out


CONST = (abi.extrapConstant,) * 6
CIRC = (abi.circular,) * 6


@pytest.mark.parametrize("method", [abi.METHOD_HLLC, abi.METHOD_HLL, abi.METHOD_XINJIN])
def test_noalloc_equals_allocating_step_twice(method):
    # first a larger grid (segmented Y/Z marches), then a smaller one on the
    # same fixed scratch array
    for shape, seed, modes, sd, sn in (((40, 72, 136), 1, CONST, 1, 3), ((12, 20, 36), 2, CIRC, -1, 4)):
        q = state(shape, seed)
        ref = step(q, method, "alloc", modes, sd, sn)
        got = step(q, method, "noalloc", modes, sd, sn)
        print(f"method {method} shape {shape}: max |diff| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.array_equal(got, ref), (method, shape)


def test_noalloc_rejects_grids_beyond_its_scratch():
    # 320^3's Y/Z marches need 82 MB of scratch (> 64 MiB): rejected before
    # any launch, the state untouched. A real 320^3 canister, so that even a
    # step that ran would stay inside its arrays.
    n = 320
    q = np.zeros((5, n, n, n))
    q[0] = 1.0
    q[1] = 1.0 / (G - 1)
    f = gc.FluidManager(q, G, bc_modes=CONST)
    gm = gc.GeometryManager((n, n, n), (1.0 / n,) * 3)
    p = gc.step_params(0.001, 1, 0, abi.METHOD_HLLC, gm, G, f.MINMASS)
    t = topo()
    rc = abi.lib().performFluidUpdate_3D_noalloc(f.arrays(), ctypes.byref(t), p)
    msg = abi.lib().imogen_last_error().decode()
    print("rejected:", rc, msg, flush=True)
    assert rc == abi.ERROR_NOMEM
    assert "allocation-free" in msg
    assert np.array_equal(f.state, q)
    f.free()
