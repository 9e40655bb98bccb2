"""Sweeps over states whose primitive differences sit near the bottom of the
double range. The precursor of a strong blast carries momenta of 1e-150 and
below (products of small velocities, shrinking by tens of decades per cell);
where two neighbouring differences A, B are ~1e-155, A B is a positive
subnormal and the Ospre denominator A^2 + A B + B^2 is subnormal too. The
reference's IEEE division then gives +-0 (gpuclass/limiter_functions.h:36-56),
and so must the sweeps, whose hardware reciprocal of a subnormal is infinite.
Parity with the oracle, NaN pattern included, for HLLC and HLL with Ospre and
van Leer, the Xin-Jin corrector's Ospre, and a Sedov blast with Ospre (the
reference's default limiter) over the iterations where its precursor first
reaches such values on the GPU's trajectory."""
import numpy as np
import pytest

from oracle import pyoracle as po

from .test_gpu_hllc import G, abi, compare, cons, gc, geom

pytestmark = pytest.mark.gpu

METHODS = {"hllc": (2, po.HLLC), "hll": (1, po.HLL), "xinjin": (3, po.XINJIN)}


def tiny_state(sh=(14, 15, 16)):
    """rho = 1, P = 1e-5 at rest, with linear momentum ramps of slope 1e-155
    per cell along every axis inside a block (so A = B ~ 1e-155 in every
    direction and component), a geometric precursor (1e-150 falling 1e-5 per
    cell) and a hot block that makes the fluxes non-trivial."""
    nz, ny, nx = sh
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rho = np.ones(sh)
    P = np.full(sh, 1e-5)
    inside = (x >= 3) & (x < 11) & (y >= 3) & (y < 11) & (z >= 3) & (z < 11)
    c = 1e-155
    p = np.zeros((3,) + sh)
    p[0] = np.where(inside, c * (1 * x + 2 * y + 3 * z - 20), 0.0)
    p[1] = np.where(inside, -c * (3 * x + 1 * y + 2 * z - 25), 0.0)
    p[2] = np.where(inside, c * (2 * x + 3 * y + 1 * z - 30), 0.0)
    decay = 1e-150 * 1e-5 ** np.arange(12)
    p[2][1, 12, :12] = decay
    p[0][12, :12, 13] = -decay
    P[11:13, 11:13, 11:13] = 1e-2
    q = cons(rho, p[0] / rho, p[1] / rho, p[2] / rho, P)
    q[2:5] = p  # the ramps exactly, not rho * (p / rho)
    return q


@pytest.mark.parametrize("case", [("hllc", 0), ("hllc", 1), ("hll", 0), ("hll", 1), ("xinjin", 0)])
@pytest.mark.parametrize("direction", [1, 2, 3])
def test_subnormal_slope_products(case, direction):
    method, limiter = case
    m, om = METHODS[method]
    q = tiny_state()
    sh = q.shape[1:]
    print(f"case {method} limiter {limiter} dir {direction}", flush=True)
    f = gc().FluidManager(q, G)
    gm = geom(sh, (0.1, 0.1, 0.1))
    gc().cudaFluidStep([0.02, 1, G, 1e-8, m, direction], f, gm, limiter=limiter)
    ref = q.copy()
    po.fluid_update_1d(ref, direction, 0.2, G, method=om, limiter=limiter)
    got = f.state
    f.free()
    assert np.all(np.isfinite(ref))
    assert np.all(np.isfinite(got))
    compare(got, ref, f"{method} limiter {limiter} dir {direction}")


def test_sedov_ospre_matches_oracle():
    """Sedov 64^3 (123 deposit cells, pressure ratio ~1e9) with HLLC and Ospre,
    three iterations through performFluidUpdate_3D: the second iteration's
    z sweep meets subnormal Ospre denominators in the precursor."""
    from gpuimogen_amd import problems
    a = abi()
    n = 64
    q0, h, _ = problems.sedov3d(n)
    f = gc().FluidManager(q0, 5 / 3, minmass=1e-8, bc_modes=(a.extrapConstant,) * 6)
    gm = gc().GeometryManager((n, n, n), h)
    ref = q0.copy()
    for it in range(3):
        tau = gc().cflTimestepFused(f, gm, a.METHOD_HLLC)
        tau_ref = po.cfl_timestep(ref, po.soundspeed(ref, 5 / 3), h, po.HLLC, (n, n, n))
        assert abs(tau - tau_ref) <= 1e-13 * tau_ref
        dt = 0.4 * tau
        for d in (1, -1):
            gc().fluidstep([f], None, None, None, [dt, 1, d, it, a.METHOD_HLLC], gm, limiter=a.LIMITER_OSPRE)
            po.fluidstep(ref, h, dt, d, it, 5 / 3, po.HLLC, po.LIM_OSPRE, modes=(po.BC_CONST,) * 6)
        got = f.state
        assert np.all(np.isfinite(got)), it
        compare(got, ref, f"sedov ospre iteration {it}")
    f.free()
