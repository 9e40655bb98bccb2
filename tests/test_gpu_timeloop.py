"""The device-side time loop (imogen_timeloop_*) against the host-driven
iteration it replaces (imogen.Simulation: imogen_cflTimestepFused per fluid,
dTime = CFL * min tau, performFluidUpdate_3D(+1), (-1) per fluid), which the
other GPU tests pin to the oracle. Same kernels, same dTime arithmetic: the
states and clocks must agree bitwise, eager or replayed from graphs. The loop
folds the CFL minimum into each iteration's last sweep where the boundaries
allow it (const with the +y/+z innermost-ghost convention, mirror, circular;
HLLC and HLL) and takes a full CFL pass otherwise (Xin-Jin) and on the first
iteration of every run call; the cases cover each."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_fluids(kind):
    from gpuimogen_amd import abi, gpuclass as gc, problems
    if kind == "sedov":
        q, h, _ = problems.sedov3d(32)
        fl = [gc.FluidManager(q, 5 / 3, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)]
        return fl, [abi.METHOD_HLLC], gc.GeometryManager((32, 32, 32), h), abi.LIMITER_VANLEER
    if kind == "twofluid":
        gas, dust, h, ciso = problems.gasdust3d(24, seed=4)
        fl = [gc.FluidManager(gas, 5 / 3, minmass=1e-8, bc_modes=(abi.circular,) * 6),
              gc.FluidManager(dust, 1.01, minmass=1e-8, bc_modes=(abi.circular,) * 6, cisothermal=ciso)]
        return fl, [abi.METHOD_HLLC, abi.METHOD_HLL], gc.GeometryManager((24, 24, 24), h), abi.LIMITER_OSPRE
    if kind == "mirror2d":
        from gpuimogen_amd import ensemble as ens
        _, st = ens.quadrant_problems(1, seed=13, nmin=40, nmax=40)
        q = np.zeros((5, 1, 40, 56))
        q[:, :, :, :40] = st[0].reshape(5, 1, 40, 40)
        q[:, :, :, 40:] = st[0].reshape(5, 1, 40, 40)[:, :, :, :16]
        fl = [gc.FluidManager(q, 1.4, minmass=1e-8, bc_modes=(abi.mirror,) * 6)]
        return fl, [abi.METHOD_HLL], gc.GeometryManager((56, 40, 1), (1 / 56, 1 / 40, 1.0)), abi.LIMITER_VANLEER
    if kind == "xinjin2d":
        from gpuimogen_amd import ensemble as ens
        _, st = ens.quadrant_problems(1, seed=11, nmin=48, nmax=48)
        fl = [gc.FluidManager(st[0].reshape(5, 1, 48, 48), 1.4, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)]
        return fl, [abi.METHOD_XINJIN], gc.GeometryManager((48, 48, 1), (1 / 48, 1 / 48, 1.0)), abi.LIMITER_OSPRE
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["sedov", "twofluid", "mirror2d", "xinjin2d"])
@pytest.mark.parametrize("graph", [False, True])
def test_device_loop_matches_host_loop(kind, graph):
    from gpuimogen_amd import imogen
    iters = 8  # more than one period of the six sweep orders
    fa, methods, gm, lim = make_fluids(kind)
    fb, _, _, _ = make_fluids(kind)
    print(f"{kind}: {iters} iterations, graph={graph}", flush=True)
    sim = imogen.Simulation(fa, methods, gm, cfl=0.4, limiter=lim)
    for _ in range(iters):
        sim.step()
    loop = imogen.DeviceLoop(fb, methods, gm, cfl=0.4, limiter=lim)
    loop.run(3, graph=graph)
    loop.run(iters - 3, graph=graph)
    t, _ = loop.clock()
    assert t == sim.time
    for a, b in zip(fa, fb):
        assert np.array_equal(a.state, b.state)
    loop.free()
    for f in fa + fb:
        f.free()


def test_device_loop_clamps_to_tmax():
    from gpuimogen_amd import imogen
    fa, methods, gm, lim = make_fluids("sedov")
    fb, _, _, _ = make_fluids("sedov")
    sim = imogen.Simulation(fa, methods, gm, cfl=0.4, limiter=lim)
    sim.step()
    tmax = sim.time * 2.5  # the third iteration is clamped
    sim.t_max = tmax
    for _ in range(3):
        sim.step()
    loop = imogen.DeviceLoop(fb, methods, gm, cfl=0.4, limiter=lim, t_max=tmax)
    loop.run(4, graph=True)
    t, dt = loop.clock()
    assert t == sim.time
    assert dt == 0.0  # the fourth iteration had no time left
    for a, b in zip(fa, fb):
        assert np.array_equal(a.state, b.state)
    loop.free()
    for f in fa + fb:
        f.free()
