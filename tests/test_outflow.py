"""Outflow boundaries (BCModeTypes outflow, gpuclass/cudaCommon.h:46; the
per-slab rules of cudaStatics.cu:486-531, applied in bounds on every face).

CPU: the oracle's outflow face against closed forms on 10-cell lines and on a
3D block -- outward flow copies the interior cell next to the face into every
ghost; inward flow keeps that cell's density and internal energy, zeroes slabs
2 and 4, sets slab 3 to rho times the x faces' rest-frame speed (1.05 on -x,
-1.198 - 0.0047 i on +x) or keeps its velocity on y/z faces; interior cells
never change. A plausible slip (the wrong side's sign, the ghost cell instead
of the interior one, total instead of internal energy kept) fails these.

GPU: setFluidBoundary and full HLLC/HLL fluid steps with outflow faces
against the oracle (bitwise for the boundaries, 1e-12 relative L1 for steps).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

G = 1.4
D = 4


def cons(rho, vx, vy, vz, P, g=G):
    q = np.zeros((5,) + np.shape(rho))
    q[0] = rho
    q[1] = P / (g - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    q[2], q[3], q[4] = rho * vx, rho * vy, rho * vz
    return q


def line_state(n, vnormal_minus, vnormal_plus, axis):
    """A 3D block of n cells along `axis` (2 across the others), with the
    normal velocity of the two interior cells next to the faces chosen."""
    sh = [2, 2, 2]
    sh[2 - (axis - 1)] = n  # numpy order (z, y, x)
    rng = np.random.default_rng(n + 10 * axis)
    rho = rng.uniform(0.5, 2.0, sh)
    P = rng.uniform(0.5, 2.0, sh)
    v = rng.uniform(-1.0, 1.0, (3,) + tuple(sh))
    ax = 2 - (axis - 1)
    idx_m = [slice(None)] * 3
    idx_m[ax] = D
    idx_p = [slice(None)] * 3
    idx_p[ax] = n - 1 - D
    v[axis - 1][tuple(idx_m)] = vnormal_minus
    v[axis - 1][tuple(idx_p)] = vnormal_plus
    return cons(rho, v[0], v[1], v[2], P)


def take(q, axis, i):
    idx = [slice(None)] * 4
    idx[1 + 2 - (axis - 1)] = i
    return q[tuple(idx)]


def eint(q):
    return q[1] - 0.5 * (q[2] ** 2 + q[3] ** 2 + q[4] ** 2) / q[0]


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_outward_flow_copies_the_interior_cell(axis):
    n = 10
    q = line_state(n, -0.5, 0.7, axis)  # both faces outward
    r = q.copy()
    po.set_fluid_boundary(r, axis, [po.BC_OUTFLOW] * 6, D)
    for k in range(D):
        assert np.array_equal(take(r, axis, k), take(q, axis, D))
        assert np.array_equal(take(r, axis, n - D + k), take(q, axis, n - 1 - D))
    assert np.array_equal(take(r, axis, slice(D, n - D)), take(q, axis, slice(D, n - D)))


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_inward_flow_rules(axis):
    n = 10
    q = line_state(n, 0.4, -0.3, axis)  # both faces inward
    r = q.copy()
    po.set_fluid_boundary(r, axis, [po.BC_OUTFLOW] * 6, D)
    for side, c in ((0, D), (1, n - 1 - D)):
        src = take(q, axis, c)
        for k in range(D):
            gi = k if side == 0 else n - D + k
            gcell = take(r, axis, gi)
            rho = src[0]
            assert np.array_equal(gcell[0], rho)
            assert np.all(gcell[2] == 0) and np.all(gcell[4] == 0)
            if axis == 1:
                v = 1.05 if side == 0 else -1.198 - 0.0047 * k
                assert np.allclose(gcell[3], rho * v, rtol=1e-15, atol=0)
            else:
                assert np.allclose(gcell[3], src[3], rtol=1e-15, atol=0)
            # internal energy of the interior cell kept
            assert np.allclose(eint(gcell), eint(src), rtol=1e-13, atol=0)
    assert np.array_equal(take(r, axis, slice(D, n - D)), take(q, axis, slice(D, n - D)))


def test_faces_decide_independently_and_zero_momentum_is_inward():
    n = 12
    q = line_state(n, -0.5, 0.0, 1)  # -x outward, +x: zero normal momentum counts as inward
    r = q.copy()
    po.set_fluid_boundary(r, 1, [po.BC_OUTFLOW] * 6, D)
    assert np.array_equal(take(r, 1, 0), take(q, 1, D))
    assert np.all(take(r, 1, n - 1)[2] == 0)
    assert np.allclose(take(r, 1, n - 1)[3], take(q, 1, n - 1 - D)[0] * (-1.198 - 0.0047 * 3), rtol=1e-15)


def test_outflow_mixed_with_const_on_the_other_face():
    n = 10
    q = line_state(n, 0.3, 0.2, 2)
    r = q.copy()
    modes = [po.BC_CONST] * 6
    modes[2] = po.BC_OUTFLOW  # -y outflow (inward there), +y const
    po.set_fluid_boundary(r, 2, modes, D)
    assert np.all(take(r, 2, 0)[2] == 0)
    for k in range(D):
        # +y const on unrotated arrays copies the innermost ghost (the y kernel's quirk)
        assert np.array_equal(take(r, 2, n - D + k), take(q, 2, n - D))


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------

def random_state(sh, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.3, 2.0, sh)
    P = rng.uniform(0.2, 2.0, sh)
    v = rng.uniform(-1.0, 1.0, (3,) + sh)
    return cons(rho, v[0], v[1], v[2], P)


@pytest.mark.gpu
@pytest.mark.parametrize("sh", [(6, 7, 8), (12, 11, 13), (16, 20, 24)])
def test_gpu_set_fluid_boundary_outflow_matches_oracle(sh):
    from gpuimogen_amd import abi, gpuclass as gc
    q = random_state(sh, seed=sum(sh))
    modes = (abi.outflow, abi.outflow, abi.outflow, abi.extrapConstant, abi.mirror, abi.outflow)
    omodes = (po.BC_OUTFLOW, po.BC_OUTFLOW, po.BC_OUTFLOW, po.BC_CONST, po.BC_MIRROR, po.BC_OUTFLOW)
    f = gc.FluidManager(q, G, bc_modes=modes)
    nz, ny, nx = sh
    gm = gc.GeometryManager((nx, ny, nz), (0.05, 0.05, 0.05))
    ref = q.copy()
    for d in (1, 2, 3):
        if sh[2 - (d - 1)] < 2 * D + 2 and d == 3:
            continue  # mirror needs 2 depth + 2 cells
        gc.setBoundaries(f, gm, d)
        po.set_fluid_boundary(ref, d, omodes)
        got = f.state
        bad = np.argwhere(got != ref)
        for b in bad[:8]:
            print("differs at", tuple(b), got[tuple(b)], ref[tuple(b)], flush=True)
        assert len(bad) == 0, (sh, d, len(bad))
    f.free()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2])
def test_gpu_fluidstep_with_outflow_matches_oracle(method):
    from gpuimogen_amd import abi, gpuclass as gc
    sh = (14, 18, 22)
    q = random_state(sh, seed=5)
    f = gc.FluidManager(q, G, bc_modes=(abi.outflow,) * 6)
    nz, ny, nx = sh
    gm = gc.GeometryManager((nx, ny, nz), (0.05, 0.05, 0.05))
    ref = q.copy()
    dt = 0.004
    for it in range(3):
        for direction in (1, -1):
            gc.fluidstep([f], None, None, None, [dt, 1, direction, it, method], gm)
            po.fluidstep(ref, gm.h, dt, direction, it, G, method, po.LIM_OSPRE, modes=(po.BC_OUTFLOW,) * 6)
    got = f.state
    f.free()
    for s in range(5):
        rel = np.abs(got[s] - ref[s]).sum() / np.abs(ref[s]).sum()
        print(f"method {method} slab {s}: rel L1 {rel:.3e}", flush=True)
        assert rel <= 1e-12
