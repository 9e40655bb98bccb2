"""The device-side time loop (imogen_timeloop_*) against the host-driven
iteration it replaces (imogen.Simulation: imogen_cflTimestepFused per fluid,
dTime = CFL * min tau, performFluidUpdate_3D(+1), (-1) per fluid), which the
other GPU tests pin to the oracle. Same kernels, same dTime arithmetic: the
states and clocks must agree bitwise, eager or replayed from graphs (X lines of
256 cells: clocks bitwise, states to an ulp, see INSTANCE_ULP). The loop
folds the CFL minimum into each iteration's last sweep where the boundaries
allow it (const with the +y/+z innermost-ghost convention, mirror, circular;
HLLC and HLL) and takes a full CFL pass otherwise (Xin-Jin) and on the first
iteration of every run call; the cases cover each. With a drag source the
loop runs imogen.m's source(1/2), fluidstep(+1), source(1), fluidstep(-1),
source(1/2) with the drag interval read from the device clock, for every
drag method that is built."""
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
    if kind == "wide":
        # X lines of 320 cells: the last sweep of an iteration may be the
        # split-half X kernel with the folded CFL minimum
        q, h, _ = problems.sedov3d(32)
        q = np.concatenate([q] * 10, axis=3)
        fl = [gc.FluidManager(q, 5 / 3, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)]
        return fl, [abi.METHOD_HLLC], gc.GeometryManager((320, 32, 32), h), abi.LIMITER_VANLEER
    if kind == "wide256":
        # X lines of 256 cells (4-cell lanes over whole lines), the last sweep
        # of some iterations with the folded CFL
        q, h, _ = problems.sedov3d(32)
        q = np.concatenate([q] * 8, axis=3)
        fl = [gc.FluidManager(q, 5 / 3, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)]
        return fl, [abi.METHOD_HLLC], gc.GeometryManager((256, 32, 32), h), abi.LIMITER_VANLEER
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


@pytest.mark.parametrize("kind", ["sedov", "wide", "wide256", "twofluid", "mirror2d", "xinjin2d"])
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
        if kind in INSTANCE_ULP:
            x, y = a.state, b.state
            rel = [float(np.abs(x[v] - y[v]).sum() / np.abs(x[v]).sum()) if np.abs(x[v]).sum() else 0.0
                   for v in range(5)]
            print(f"{kind}: relative L1 difference per variable {rel}", flush=True)
            assert max(rel) <= 1e-14, rel
        else:
            assert np.array_equal(a.state, b.state)
    loop.free()
    for f in fa + fb:
        f.free()


# Lines of 256 cells are swept with 4 cells per lane over whole lines; the
# folded-CFL form of that kernel (the device loop's last sweep) is a template
# instance of its own whose FMA contraction differs from the plain one by an
# ulp at some cells, so there the loops agree to rounding, not bitwise.
INSTANCE_ULP = {"wide256"}


def make_gasdust(n=16, seed=6, grain_mass=1.0, isothermal_dust=False):
    from gpuimogen_amd import abi, gpuclass as gc, learned_drag as L, problems
    gas, dust, h, ciso = problems.gasdust3d(n, seed=seed)
    tg, td = L.cold_h2(), L.iron_dust()
    td.m *= grain_mass
    fl = [gc.FluidManager(gas, tg.gamma, minmass=1e-8, bc_modes=(abi.circular,) * 6, thermoDetails=tg),
          gc.FluidManager(dust, td.gamma, minmass=1e-8, bc_modes=(abi.circular,) * 6, thermoDetails=td,
                          cisothermal=ciso if isothermal_dust else -1.0)]
    methods = [abi.METHOD_HLLC, abi.METHOD_HLL if isothermal_dust else abi.METHOD_HLLC]
    return fl, methods, gc.GeometryManager((n, n, n), h)


@pytest.mark.parametrize("method", ["EMP", "RK4", "ETDRK1", "LOGTRAP2", "LOGTRAP3", "EXACT", "LEARNED"])
@pytest.mark.parametrize("graph", [False, True])
def test_device_loop_with_drag_matches_host_loop(method, graph):
    """source(1/2), fluidstep(+1), source(1), fluidstep(-1), source(1/2) with
    the drag interval read from the device clock: the same kernels and the
    same dt as the host loop's src0001 calls, so bitwise the same states."""
    from gpuimogen_amd import abi, imogen, learned_drag as L
    drag = getattr(abi, "DRAG_" + method)
    if drag == abi.DRAG_LEARNED:
        params, norm, _ = L.load_model()
        L.set_model(params, norm)
    iters = 7
    # EMP and RK4 rebuild the dust energy as its kinetic energy + 1e-4 rho
    # (cukern_applyFinalDeltaV: pressureless dust), which an HLLC dust sweep
    # turns negative; they are run with the isothermal HLL dust. At 16^3 the
    # drag rate reaches k dt ~ 3, beyond these explicit integrators: they get
    # a 20x heavier grain (20x smaller rate). The oracle shows the same.
    explicit = drag in (abi.DRAG_EMP, abi.DRAG_RK4)
    mass = 20.0 if explicit else 1.0
    fa, methods, gm = make_gasdust(grain_mass=mass, isothermal_dust=explicit)
    fb, _, _ = make_gasdust(grain_mass=mass, isothermal_dust=explicit)
    print(f"gas+dust 16^3, drag {method}: {iters} iterations, graph={graph}", flush=True)
    sim = imogen.Simulation(fa, methods, gm, cfl=0.4, drag_method=drag)
    for _ in range(iters):
        sim.step()
    L.learned_stats()
    loop = imogen.DeviceLoop(fb, methods, gm, cfl=0.4, drag_method=drag)
    loop.run(2, graph=graph)
    loop.run(iters - 2, graph=graph)
    t, _ = loop.clock()
    assert t == sim.time
    for a, b in zip(fa, fb):
        assert np.all(np.isfinite(a.state))
        assert np.array_equal(a.state, b.state)
    if drag == abi.DRAG_LEARNED:
        # every source call is two drag calls over all cells, replays included;
        # some cells are rejected (the device loop takes a source's two calls in
        # one pass, and a cell whose first call is rejected meets its second
        # call in a second pass over a list: both paths are exercised here)
        cells, nfall = L.learned_stats()
        print(f"learned cells {cells}, fallbacks {nfall}", flush=True)
        assert cells == 6 * iters * 16 ** 3
        assert nfall > 0
    loop.free()
    for f in fa + fb:
        f.free()


def test_device_loop_drag_needs_two_fluids():
    from gpuimogen_amd import abi, imogen
    fl, methods, gm, lim = make_fluids("sedov")
    with pytest.raises(ValueError):
        imogen.DeviceLoop(fl, methods, gm, drag_method=abi.DRAG_LOGTRAP2)
    fa, methods, gm = make_gasdust(8)
    with pytest.raises(abi.ImogenError):
        imogen.DeviceLoop(fa, methods, gm, drag_method=abi.DRAG_ETDRK2)
    for f in fl + fa:
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
