"""Index exchanges of device arrays (flipArrayIndices, the cudaArrayRotateB MEX)
and the reference's MATLAB-side fluid step that relies on them (flux/flux.m:
rotate so the swept direction is memory x, cudaFluidStep, cudaStatics per
array, halo exchange, rotate on).

Pins: every exchange code moves the data exactly as the corresponding numpy
transpose and rewrites dims / permtag accordingly; flux.m's rotated sequence
gives bitwise the state of performFluidUpdate_3D, which never rotates. The
latter fixes the boundary conventions inside the fluid step: the reference sets
each sweep's boundaries with that direction in memory x (flux.cu:99-107), so
its memory-x kernels act along every direction."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AXES = {2: (1, 0, 2), 3: (2, 1, 0), 4: (0, 2, 1), 5: (1, 2, 0), 6: (2, 0, 1)}


def gc():
    from gpuimogen_amd import gpuclass
    return gpuclass


def abi():
    from gpuimogen_amd import abi as a
    return a


def expected_rotation(host, code):
    """host is (nz, ny, nx); returns the rotated array in the same
    (slowest..fastest) layout and its memory dims."""
    a_xyz = host.transpose(2, 1, 0)
    b = np.transpose(a_xyz, AXES[code])
    return np.ascontiguousarray(b.transpose(2, 1, 0)), b.shape


@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 7, 9), (4, 1, 6)])
@pytest.mark.parametrize("code", [2, 3, 4, 5, 6])
def test_rotation_is_the_transpose(shape, code):
    rng = np.random.default_rng(code)
    host = rng.standard_normal(shape)
    print(f"rotate {shape} code {code}", flush=True)
    g = gc().GPU_Type(host)
    gc().cudaArrayRotateB(g, code)
    want, dims = expected_rotation(host, code)
    assert g.dims == dims
    perm = (ctypes.c_int * 3)(*[AXES[code][a] + 1 for a in range(3)])
    assert int(g.tag[8]) == abi().lib().MGA_numsToPermtag(perm)
    got = g.array
    assert got.shape == want.shape and np.array_equal(got, want)
    g.free()


@pytest.mark.parametrize("codes", [(2, 2), (3, 3), (4, 4), (5, 6), (6, 5), (5, 5, 5), (6, 6, 6)])
def test_rotations_compose_to_identity(codes):
    host = np.random.default_rng(1).standard_normal((6, 5, 8))
    g = gc().GPU_Type(host)
    for c in codes:
        gc().cudaArrayRotateB(g, c)
    assert g.dims == (8, 5, 6) and int(g.tag[8]) == 1
    assert np.array_equal(g.array, host)
    g.free()


def test_rotation_into_new_arrays():
    a = abi()
    host = np.random.default_rng(2).standard_normal((4, 6, 10))
    g = gc().GPU_Type(host)
    m = g.mgarray()
    nu = (a.MGArray * 1)()
    ptr = ctypes.cast(nu, ctypes.POINTER(a.MGArray))
    a.call("flipArrayIndices", ctypes.byref(m), ctypes.byref(ptr), 1, 5, None, None)
    assert tuple(nu[0].dim) == (6, 4, 10) and nu[0].numSlabs == 1
    out = np.empty(4 * 6 * 10)
    a.call("MGA_downloadArrayToCPU", ctypes.byref(nu[0]),
           ctypes.byref(ctypes.cast(out.ctypes.data, ctypes.POINTER(ctypes.c_double))), 0)
    want, _ = expected_rotation(host, 5)
    assert np.array_equal(out.reshape(want.shape), want)
    assert np.array_equal(g.array, host)  # the source is untouched
    a.call("MGA_delete", ctypes.byref(nu[0]))
    g.free()


def smooth_state(shape, seed):
    """A state whose boundary cells differ, so every ghost-cell convention
    shows in the result."""
    from gpuimogen_amd import problems
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rho = 1 + 0.3 * np.sin(0.7 * x + 0.3) * np.cos(0.5 * y) + 0.2 * np.cos(0.9 * z + 1)
    v = [0.4 * np.sin(0.4 * x + k) * np.cos(0.6 * y - k) + 0.1 * rng.standard_normal(shape) for k in range(3)]
    P = 1 + 0.25 * np.cos(0.3 * x + 0.8 * y - 0.5 * z)
    return problems.conserved(rho, v[0], v[1], v[2], P, 5 / 3)


@pytest.mark.parametrize("bc", ["const", "mirror"])
@pytest.mark.parametrize("order", [1, -1])
def test_flux_m_rotated_path_equals_fluidstep(bc, order):
    a = abi()
    shape = (14, 12, 16)
    q = smooth_state(shape, 3)
    mode = a.extrapConstant if bc == "const" else a.mirror
    gm = gc().GeometryManager((16, 12, 14), (1 / 16, 1 / 12, 1 / 14))
    fa = gc().FluidManager(q, 5 / 3, bc_modes=(mode,) * 6)
    fb = gc().FluidManager(q, 5 / 3, bc_modes=(mode,) * 6)
    dt = 0.3 * gc().cflTimestepFused(fa, gm, a.METHOD_HLLC)
    for it in range(4):
        print(f"flux.m path, {bc} faces, order {order}, iteration {it}", flush=True)
        gc().flux([fa], gm, order, it, dt, a.METHOD_HLLC)
        gc().fluidstep([fb], None, None, None, [dt, 1, order, it, a.METHOD_HLLC], gm)
        for g in [fa.mass, fa.ener] + fa.mom:
            assert int(g.tag[8]) == 1 and g.dims == (16, 12, 14)  # flux.m ends in XYZ order
        assert np.array_equal(fa.state, fb.state)
    fa.free()
    fb.free()


def test_rotated_fluid_through_the_entry_points():
    """A fluid whose arrays are rotated (as flux.m leaves them between
    sweeps): its CFL limit and its 1D updates in physical directions are
    those of the unrotated fluid. (Periodic faces: a rotated fluid's
    boundaries use the kernels of the memory axis holding each direction,
    cudaStatics.cu:230, which differ from the unrotated fluid's for const and
    z-mirror faces; the flux.m test above covers those.)"""
    a = abi()
    shape = (10, 8, 12)
    q = smooth_state(shape, 4)
    gm = gc().GeometryManager((12, 8, 10), (1 / 12, 1 / 8, 1 / 10))
    fa = gc().FluidManager(q, 5 / 3, bc_modes=(a.circular,) * 6)
    fb = gc().FluidManager(q, 5 / 3, bc_modes=(a.circular,) * 6)
    tau = gc().cflTimestepFused(fa, gm, a.METHOD_HLLC)
    gc().xchgIndices(fa, 6)  # XYZ -> ZXY: z is now memory x
    assert gc().cflTimestepFused(fa, gm, a.METHOD_HLLC) == tau
    for d in (3, 1, 2):
        gc().relaxingFluid(fa, gm, d, 0.3 * tau, a.METHOD_HLLC)
        gc().relaxingFluid(fb, gm, d, 0.3 * tau, a.METHOD_HLLC)
    gc().xchgIndices(fa, 5)  # back to XYZ
    assert np.array_equal(fa.state, fb.state)
    fa.free()
    fb.free()
