"""Freebalance boundaries and static cells (BCModeTypes freebalance,
gpuclass/cudaCommon.h:48; setFreeBalanceCondition and
cukern_SetFreeBalance{X,Y,Z}, cudaStatics.cu:700-1003; setArrayStaticCells and
cukern_applySpecial_fade, :288-350, 1251-1281).

CPU: the oracle's freebalance face against physics it must reproduce --
a uniform potential gives constant extrapolation of (rho, v, P); a linear
potential gives the isothermal atmosphere rho ~ exp(-phi rho_c / P_c) at the
reference cell's temperature; on x faces an azimuthal velocity follows a Kepler
curve v ~ r^-1/2 in the inertial frame and its centrifugal potential enters the
balance; tenuous cells (rho < 1e-7) ignore the potential and drop an inflowing
z velocity; interior cells never change. The static-cell fade against its
closed forms. A plausible slip (a sign of the potential difference, the ghost
instead of the reference cell, a missing frame velocity, the wrong Kepler
exponent) fails these.

GPU: setFreeBalanceCondition, applyStaticCells and setFluidBoundary with a
freebalance face and static cells against the oracle (exp in the ghost density:
1e-14 relative; the fade bitwise)."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

G = 5.0 / 3.0
D = 4


def cons(rho, vx, vy, vz, P, g=G):
    q = np.zeros((5,) + np.shape(rho))
    q[0] = rho
    q[1] = P / (g - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    q[2], q[3], q[4] = rho * vx, rho * vy, rho * vz
    return q


def block(seed, sh=(6, 7, 12), rho=(0.5, 2.0)):
    rng = np.random.default_rng(seed)
    r = rng.uniform(*rho, sh)
    P = rng.uniform(0.5, 2.0, sh)
    v = rng.uniform(-1.0, 1.0, (3,) + sh)
    return cons(r, v[0], v[1], v[2], P)


def axis_index(direction):
    return 3 - direction  # numpy axis of (nz, ny, nx) holding direction 1/2/3


def take(a, direction, i):
    idx = [slice(None)] * a.ndim
    idx[a.ndim - direction] = i
    return a[tuple(idx)]


def coord(sh, direction):
    n = sh[axis_index(direction)]
    shape = [1, 1, 1]
    shape[axis_index(direction)] = n
    return np.arange(n, dtype=np.float64).reshape(shape) * np.ones(sh)


def prim(q):
    rho = q[0]
    v = q[2:5] / rho
    P = (G - 1) * (q[1] - 0.5 * rho * (v ** 2).sum(0))
    return rho, v, P


@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("side", [0, 1])
def test_uniform_potential_extrapolates_constant(direction, side):
    q = block(10 * direction + side)
    if direction == 1:
        q[3] = 0.0  # no azimuthal velocity: no centrifugal term on x faces
        q[1] = cons(q[0], q[2] / q[0], 0 * q[0], q[4] / q[0], prim(q)[2])[1]
    phi = np.full(q.shape[1:], 0.37)
    ref = q.copy()
    po.freebalance_face(q, phi, direction, side, G, 1.0, 0.1)  # radius 1 at the grid's first cell
    n = q.shape[1 + axis_index(direction)]
    c = n - 1 - D if side else D
    ghosts = range(n - D, n) if side else range(D)
    for gi in ghosts:
        np.testing.assert_allclose(take(q, direction, gi), take(ref, direction, c), rtol=1e-14, atol=0)
    inner = range(0, n - D) if side else range(D, n)
    for i in inner:
        assert np.array_equal(take(q, direction, i), take(ref, direction, i))


@pytest.mark.parametrize("direction", [2, 3])
@pytest.mark.parametrize("side", [0, 1])
def test_linear_potential_gives_isothermal_atmosphere(direction, side):
    q = block(20 * direction + side)
    g = 0.3
    phi = g * coord(q.shape[1:], direction)
    ref = q.copy()
    po.freebalance_face(q, phi, direction, side, G)
    n = q.shape[1 + axis_index(direction)]
    c = n - 1 - D if side else D
    rc, vc, Pc = prim(take(ref, direction, c))
    for gi in (range(n - D, n) if side else range(D)):
        rg, vg, Pg = prim(take(q, direction, gi))
        # hydrostatic isothermal gas: dP = -rho dphi with P/rho fixed
        expect = rc * np.exp(-g * (gi - c) * rc / Pc)
        np.testing.assert_allclose(rg, expect, rtol=1e-13)
        np.testing.assert_allclose(Pg / rg, Pc / rc, rtol=1e-13)
        np.testing.assert_allclose(vg, vc, rtol=1e-13)


@pytest.mark.parametrize("side", [0, 1])
def test_x_faces_follow_a_kepler_curve(side):
    """Direction 1 is radial in the disk setups: r = Rinner + i h0, frame
    rotating at Omega. The inertial azimuthal velocity w = vy + Omega r_c is
    extrapolated as w (r_c / r)^1/2, and the gas balances the potential plus
    the centrifugal potential of that curve, -w^2 (1 - r_c / r)."""
    q = block(40 + side)
    rinner, h0, om = 2.0, 0.05, 0.7
    nx = q.shape[3]
    phi = -1.0 / (rinner + h0 * coord(q.shape[1:], 1))  # point-mass potential
    ref = q.copy()
    po.freebalance_face(q, phi, 1, side, G, rinner, h0, om)
    c = nx - 1 - D if side else D
    r_c = rinner + c * h0
    rc, vc, Pc = prim(take(ref, 1, c))
    w = vc[1] + om * r_c
    for gi in (range(nx - D, nx) if side else range(D)):
        r_g = rinner + gi * h0
        rg, vg, Pg = prim(take(q, 1, gi))
        np.testing.assert_allclose(vg[1] + om * r_g, w * np.sqrt(r_c / r_g), rtol=1e-13)
        np.testing.assert_allclose(vg[0], vc[0], rtol=1e-14)
        np.testing.assert_allclose(vg[2], vc[2], rtol=1e-14)
        dphi = take(phi, 1, gi) - take(phi, 1, c) - w ** 2 * (1 - r_c / r_g)
        np.testing.assert_allclose(rg, rc * np.exp(-dphi * rc / Pc), rtol=1e-12)
        np.testing.assert_allclose(Pg / rg, Pc / rc, rtol=1e-13)


@pytest.mark.parametrize("side", [0, 1])
def test_tenuous_cells_ignore_the_potential_and_stop_inflow(side):
    q = block(50 + side, rho=(1e-9, 5e-8))
    # a z velocity along +z on the minus face (-z on the plus face) on half the
    # lines (x < 6), the opposite sign on the rest; the reference clamps the z
    # component on every face (the disk setups' vertical direction)
    inward = 1.0 if side == 0 else -1.0
    q[4, :, :, :6] = q[0, :, :, :6] * 0.4 * inward
    q[4, :, :, 6:] = -q[0, :, :, 6:] * 0.4 * inward
    phi = 5.0 * coord(q.shape[1:], 2)
    ref = q.copy()
    po.freebalance_face(q, phi, 2, side, G)
    n = q.shape[2]
    c = n - 1 - D if side else D
    rc, vc, Pc = prim(take(ref, 2, c))
    vz = np.where(vc[2] * (1.0 if side else -1.0) < 0, 0.0, vc[2])
    for gi in (range(n - D, n) if side else range(D)):
        rg, vg, Pg = prim(take(q, 2, gi))
        np.testing.assert_allclose(rg, rc, rtol=1e-15)
        np.testing.assert_allclose(vg[2], vz, rtol=1e-14, atol=1e-300)
        assert np.all(vg[2][:, :6] == 0.0) and np.all(vg[2][:, 6:] != 0.0)


def test_static_cell_fade_closed_forms():
    rng = np.random.default_rng(3)
    phi = rng.uniform(-1, 1, (4, 5, 6))
    ref = phi.copy()
    phi[1, 1, 1] = np.nan
    idx = np.array([7, 20, 31, 99, 119])
    f0 = np.array([2.0, -1.5, 0.25, 4.0, 1.0])
    c = np.array([1.0, 0.25, 0.0, 0.5, 0.5])
    flat = ref.reshape(-1).copy()
    nanat = 1 * 30 + 1 * 6 + 1  # phi[1, 1, 1]
    idx[4] = nanat
    po.static_cells(phi, np.stack([idx, f0, c]).astype(np.float64))
    got = phi.reshape(-1)
    assert got[7] == 2.0                           # c = 1: the stored value
    assert got[20] == -1.5 * 0.25 + 0.75 * flat[20]  # fade toward f0
    assert got[31] == flat[31]                     # c = 0: unchanged (f0 * 0 + 1 * phi)
    assert got[99] == 4.0 * 0.5 + 0.5 * flat[99]
    assert got[nanat] == 1.0 * 0.5                 # NaN: f0 c
    others = np.setdiff1d(np.arange(flat.size), idx)
    assert np.array_equal(got[others], flat[others])


def test_abi_null_arguments_fail_without_device_work():
    from gpuimogen_amd import abi
    L = abi.lib()
    assert L.setFreeBalanceCondition(None, None, 0, 1, None, 1.4) == abi.ERROR_NULL_POINTER
    assert L.applyStaticCells(None, None, None) == abi.ERROR_NULL_POINTER


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------

def _gpu():
    from gpuimogen_amd import abi, gpuclass as gc
    return abi, gc


def geometry(gc, q, rinner=0.0, om=0.0, h=(0.05, 0.05, 0.05)):
    gm = gc.GeometryManager(q.shape[:0:-1], h)
    g = gm.params()
    g.Rinner = rinner
    g.frameOmega = om
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("tenuous", [False, True])
def test_set_freebalance_condition_matches_oracle(direction, side, tenuous):
    abi, gc = _gpu()
    q = block(60 + 7 * direction + side, sh=(9, 10, 11), rho=(1e-9, 5e-8) if tenuous else (0.5, 2.0))
    rinner, om = 1.5, 0.4
    phi = -1.0 / (rinner + 0.05 * coord(q.shape[1:], 1)) + 0.2 * coord(q.shape[1:], 3)
    print(f"freebalance dir {direction} side {side} tenuous {tenuous}", flush=True)
    f = gc.FluidManager(q, G)
    pot = gc.GPU_Type(phi)
    arr = f.arrays()
    g = geometry(gc, q, rinner, om)
    pm = pot.mgarray()
    abi.call("setFreeBalanceCondition", arr, ctypes.byref(g), side, direction, ctypes.byref(pm), G)
    got = f.state
    ref = q.copy()
    po.freebalance_face(ref, phi, direction, side, G, rinner, 0.05, om)
    n = q.shape[1 + axis_index(direction)]
    ghosts = list(range(n - D, n)) if side else list(range(D))
    inner = [i for i in range(n) if i not in ghosts]
    for i in inner:
        assert np.array_equal(take(got, direction, i), take(q, direction, i))
    for i in ghosts:
        a, b = take(got, direction, i), take(ref, direction, i)
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=0)
    f.free()
    pot.free()


@pytest.mark.gpu
def test_apply_static_cells_matches_oracle():
    abi, gc = _gpu()
    rng = np.random.default_rng(9)
    phi = rng.uniform(-1, 1, (5, 6, 7))
    n = 40
    idx = rng.choice(phi.size, n, replace=False).astype(np.float64)
    f0 = rng.uniform(-2, 2, n)
    c = np.where(rng.uniform(size=n) < 0.3, 1.0, rng.uniform(0, 1, n))
    # the table holds this orientation's entries after 5 entries of another
    table = np.zeros((3, n + 5))
    table[:, 5:] = np.stack([idx, f0, c])
    comp = (ctypes.c_double * (12 * abi.MAX_GPUS_USED))()
    comp[0], comp[1] = 5.0, float(n)  # orientation 0 (XYZ): offset, count
    a = gc.GPU_Type(phi)
    t = gc.GPU_Type(table)  # 3 columns of n + 5 rows: dim[0] = n + 5
    am, tm = a.mgarray(), t.mgarray()
    assert tm.dim[0] == n + 5
    abi.call("applyStaticCells", ctypes.byref(am), ctypes.byref(tm), comp)
    got = a.array
    ref = phi.copy()
    po.static_cells(ref, table[:, 5:])
    assert np.array_equal(got, ref)
    a.free()
    t.free()


@pytest.mark.gpu
def test_set_fluid_boundary_with_freebalance_face_and_statics():
    """-x freebalance, +x const; static cells on the density and energy
    arrays: each face runs after every array's static cells (the reference
    applies them once per face), bitwise for the const face and the fade,
    1e-14 for the freebalance densities."""
    abi, gc = _gpu()
    q = block(77, sh=(5, 6, 13))
    modes = (abi.freebalance, abi.extrapConstant) + (abi.extrapConstant,) * 4
    phi = 0.1 * coord(q.shape[1:], 1) ** 2
    rng = np.random.default_rng(5)
    n = 12
    tabs, keep = [], []
    f = gc.FluidManager(q, G, bc_modes=modes)
    pot = gc.GPU_Type(phi)
    pm = pot.mgarray()
    arr = f.arrays()
    bds = []
    for k in range(5):
        bd = abi.ImogenBoundaryData()
        if k == 0:
            bd.gravpot = ctypes.pointer(pm)
            bd.gamma = G
        if k in (0, 1):
            idx = rng.choice(np.arange(4, 9), n).astype(np.float64) + 13 * rng.integers(0, 30, n)
            idx = np.unique(idx)[:n]
            m = idx.size
            tab = np.stack([idx, rng.uniform(0.5, 2, m), rng.choice([1.0, 0.5], m)])
            t = gc.GPU_Type(tab)
            tm = t.mgarray()
            keep += [t, tm]
            bd.statics = ctypes.pointer(tm)
            bd.compOffset[0], bd.compOffset[1] = 0.0, float(m)
            tabs.append(tab)
        else:
            tabs.append(None)
        bds.append(bd)
        arr[k].boundaryConditions.externalData = ctypes.cast(ctypes.pointer(bd), ctypes.c_void_p)
    g = geometry(gc, q)
    abi.call("setFluidBoundary", arr, ctypes.byref(g), 1)
    got = f.state
    ref = q.copy()
    for side in (0, 1):
        for k in (0, 1):
            po.static_cells(ref[k], tabs[k])
        if side == 0:
            po.freebalance_face(ref, phi, 1, 0, G, 0.0, 0.05, 0.0)
        else:
            for s in range(5):
                ref[s][:, :, 13 - D:] = ref[s][:, :, 13 - D - 1:13 - D]
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)
    assert np.array_equal(got[:, :, :, D:], ref[:, :, :, D:])
    f.free()
    pot.free()
    for t in keep[::2]:
        t.free()
