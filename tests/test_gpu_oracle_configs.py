"""Every configuration's GPU path against the CPU oracle directly (not through
another GPU path):

  device loop  Sedov-Taylor 32^3 (HLLC, van Leer, const boundaries, CFL
               folded into the last sweep) and gas + dust 16^3 (both HLLC,
               periodic) with LogTrap2 and with the adaptive integrator, and
               with the learned drag against the oracle's adaptive integrator
  ensemble     six 2D four-quadrant problems of 64^2..200^2, three with dust
               (LogTrap2; learned vs the oracle's adaptive integrator), each
               problem with its own CFL time step
  Sod          Xin-Jin at the configuration's N = 1024, run to t = 0.2 by the
               device loop's t_max clamp; also against the exact solution
  Orszag-Tang  the configuration's 1024^2 periodic initial state on
               [0, 2 pi]^2 with gamma = 5/3, rho = gamma^2, P = gamma,
               v = (-sin y, sin x), Xin-Jin at CFL 0.8 -- a hydrodynamic
               proxy: the magnetic field is left out (the reference's MHD
               path is not built, and neither is it here)

Bars: 1e-12 relative L1 per conserved variable for the same arithmetic; 1e-10
where the adaptive integrator's step controller can take a different step
sequence on rounding differences; the learned drag within its residual
tolerance's reach (1e-6) of the oracle's adaptive integrator.
"""
import numpy as np
import pytest

from oracle import exact_riemann as er
from oracle import pyoracle as po

from .oracle_loop import OracleFluid, oracle_iterations, rel_l1
from .test_oracle_drag import gas_dust_thermo

pytestmark = pytest.mark.gpu


def _thermo_to_abi(t):
    from gpuimogen_amd import abi
    return abi.ThermoDetails(*[getattr(t, f) for f, _ in t._fields_])


def _check(got, ref, bar, what):
    err = rel_l1(got, ref)
    print(f"{what}: rel L1 per slab {np.array2string(err, precision=2)}", flush=True)
    assert np.all(np.isfinite(got))
    assert err.max() <= bar, (what, err)


# ---------------------------------------------------------------- device loop
def test_device_loop_sedov_matches_oracle():
    from gpuimogen_amd import abi, gpuclass as gc, imogen, problems
    q, h, _ = problems.sedov3d(32)
    iters = 8
    f = gc.FluidManager(q, 5 / 3, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)
    loop = imogen.DeviceLoop([f], [abi.METHOD_HLLC], gc.GeometryManager((32, 32, 32), h), cfl=0.4,
                             limiter=abi.LIMITER_VANLEER)
    print(f"device loop Sedov 32^3: {iters} iterations", flush=True)
    loop.run(iters, graph=True)
    t, _ = loop.clock()
    got = f.state
    loop.free()
    f.free()
    ref = OracleFluid(q, 5 / 3, po.HLLC, (po.BC_CONST,) * 6, po.LIM_VANLEER)
    tref = oracle_iterations([ref], h, iters)
    assert t == pytest.approx(tref, rel=1e-13)
    _check(got, ref.q, 1e-12, "Sedov 32^3 device loop vs oracle")


def _gasdust(n=16, seed=6):
    from gpuimogen_amd import problems
    gas, dust, h, _ = problems.gasdust3d(n, seed=seed)
    return gas, dust, h


@pytest.mark.parametrize("method,bar", [("LOGTRAP2", 1e-12), ("EXACT", 1e-10)])
def test_device_loop_gasdust_matches_oracle(method, bar):
    from gpuimogen_amd import abi, gpuclass as gc, imogen
    gas, dust, h = _gasdust()
    tg, td = gas_dust_thermo()
    drag = getattr(abi, "DRAG_" + method)
    iters = 5
    modes = (abi.circular,) * 6
    fg = gc.FluidManager(gas, tg.gamma, minmass=1e-8, bc_modes=modes, thermoDetails=_thermo_to_abi(tg))
    fd = gc.FluidManager(dust, td.gamma, minmass=1e-8, bc_modes=modes, thermoDetails=_thermo_to_abi(td))
    loop = imogen.DeviceLoop([fg, fd], [abi.METHOD_HLLC] * 2, gc.GeometryManager((16, 16, 16), h), cfl=0.4,
                             drag_method=drag)
    print(f"device loop gas+dust 16^3 drag {method}: {iters} iterations", flush=True)
    loop.run(iters, graph=True)
    t, _ = loop.clock()
    got = fg.state, fd.state
    loop.free()
    fg.free()
    fd.free()
    rg = OracleFluid(gas, tg.gamma, po.HLLC, (po.BC_CIRCULAR,) * 6, thermo=tg)
    rd = OracleFluid(dust, td.gamma, po.HLLC, (po.BC_CIRCULAR,) * 6, thermo=td)
    tref = oracle_iterations([rg, rd], h, iters, drag=drag)
    assert t == pytest.approx(tref, rel=1e-12)
    _check(got[0], rg.q, bar, f"gas, drag {method}")
    _check(got[1], rd.q, bar, f"dust, drag {method}")


def test_device_loop_learned_drag_tracks_oracle_exact_drag():
    from gpuimogen_amd import abi, gpuclass as gc, imogen, learned_drag as L
    params, norm, _ = L.load_model()
    L.set_model(params, norm)
    gas, dust, h = _gasdust(seed=7)
    tg, td = gas_dust_thermo()
    iters = 5
    modes = (abi.circular,) * 6
    fg = gc.FluidManager(gas, tg.gamma, minmass=1e-8, bc_modes=modes, thermoDetails=_thermo_to_abi(tg))
    fd = gc.FluidManager(dust, td.gamma, minmass=1e-8, bc_modes=modes, thermoDetails=_thermo_to_abi(td))
    loop = imogen.DeviceLoop([fg, fd], [abi.METHOD_HLLC] * 2, gc.GeometryManager((16, 16, 16), h), cfl=0.4,
                             drag_method=abi.DRAG_LEARNED)
    print(f"device loop gas+dust 16^3 learned drag: {iters} iterations", flush=True)
    loop.run(iters, graph=True)
    got = fg.state, fd.state
    loop.free()
    fg.free()
    fd.free()
    rg = OracleFluid(gas, tg.gamma, po.HLLC, (po.BC_CIRCULAR,) * 6, thermo=tg)
    rd = OracleFluid(dust, td.gamma, po.HLLC, (po.BC_CIRCULAR,) * 6, thermo=td)
    oracle_iterations([rg, rd], h, iters, drag=po.DRAG_EXACT)
    _check(got[0], rg.q, 1e-6, "gas, learned drag vs oracle exact drag")
    _check(got[1], rd.q, 1e-6, "dust, learned drag vs oracle exact drag")


# ---------------------------------------------------------------- ensemble
ENS_SIZES = [64, 90, 128, 77, 200, 150]


@pytest.mark.parametrize("drag", ["LOGTRAP2", "LEARNED"])
def test_ensemble_matches_oracle_per_problem(drag):
    from gpuimogen_amd import abi, ensemble as ens, learned_drag as L
    if drag == "LEARNED":
        params, norm, _ = L.load_model()
        L.set_model(params, norm)
    tg, td = gas_dust_thermo()
    tg.gamma = 1.4  # the quadrant problems' gas
    rng = np.random.default_rng(31)
    gas, dust = [], {}
    for p, n in enumerate(ENS_SIZES):
        _, st = ens.quadrant_problems(1, seed=int(rng.integers(1 << 30)), nmin=n, nmax=n)
        gas.append(st[0])
        if p % 2:
            dust[p] = ens.quadrant_dust(st[0], seed=200 + p)
    sizes = np.array(ENS_SIZES)
    e = ens.Ensemble(sizes, sizes, 1.0 / sizes, 1.0 / sizes)
    for p, s in enumerate(gas):
        e.set_state(p, s)
    e.set_dust([p in dust for p in range(len(sizes))], _thermo_to_abi(tg), _thermo_to_abi(td), 1.01,
               getattr(abi, "DRAG_" + drag))
    for p, s in dust.items():
        e.set_dust_state(p, s)
    iters = 3
    print(f"ensemble of {len(sizes)} problems ({drag} drag on {len(dust)}): {iters} iterations", flush=True)
    e.run(iters, graph=True)
    times = e.times()
    bar = 1e-12 if drag == "LOGTRAP2" else 1e-6
    for p, n in enumerate(ENS_SIZES):
        h = (1.0 / n, 1.0 / n, 1.0)
        fl = [OracleFluid(gas[p].reshape(5, 1, n, n), 1.4, po.HLLC, (po.BC_CONST,) * 6, thermo=tg)]
        if p in dust:
            fl.append(OracleFluid(dust[p].reshape(5, 1, n, n), 1.01, po.HLLC, (po.BC_CONST,) * 6, thermo=td))
        oracle_drag = None
        if p in dust:
            oracle_drag = po.DRAG_LOGTRAP2 if drag == "LOGTRAP2" else po.DRAG_EXACT
        t = oracle_iterations(fl, h, iters, drag=oracle_drag)
        assert times[p] == pytest.approx(t, rel=1e-12 if drag == "LOGTRAP2" else 1e-9)
        _check(e.state(p), fl[0].q.reshape(5, n, n), bar, f"problem {p} ({n}^2) gas")
        if p in dust:
            _check(e.dust_state(p), fl[1].q.reshape(5, n, n), bar, f"problem {p} ({n}^2) dust")
    e.free()


# ---------------------------------------------------------------- Sod N = 1024
def test_sod_1024_xinjin_to_t02_matches_oracle_and_exact_solution():
    from gpuimogen_amd import abi, gpuclass as gc, imogen, problems
    q, h = problems.sod1d(1024)
    f = gc.FluidManager(q, 1.4, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)
    loop = imogen.DeviceLoop([f], [abi.METHOD_XINJIN], gc.GeometryManager((1024, 1, 1), h), cfl=0.4, t_max=0.2)
    ref = OracleFluid(q, 1.4, po.XINJIN, (po.BC_CONST,) * 6)
    # iterations until the oracle's clock reaches t_max (the last one clamped)
    t, iters = 0.0, 0
    while t < 0.2 * (1 - 1e-14):
        t = oracle_iterations([ref], h, 1, first_iteration=iters, t0=t, t_max=0.2)
        iters += 1
    print(f"Sod 1024 Xin-Jin: {iters} iterations to t = {t}", flush=True)
    loop.run(iters, graph=True)
    tg, _ = loop.clock()
    got = f.state
    loop.free()
    f.free()
    assert tg == pytest.approx(0.2, rel=1e-14)
    # quiescent slabs (py, pz) are zero on both sides; the measure is per slab
    d = np.abs(got - ref.q).reshape(5, -1).sum(1)
    scale = np.maximum(np.abs(ref.q).reshape(5, -1).sum(1), np.abs(ref.q).sum() * 1e-3)
    print(f"Sod 1024 GPU vs oracle: rel L1 {d / scale}", flush=True)
    assert np.all(d / scale <= 1e-12)
    x = (np.arange(1024) + 0.5) / 1024
    ex = er.cell_averages(x, 1.0 / 1024, 0.5, 0.2, (1, 0, 1), (0.125, 0, 0.1), 1.4)
    l1 = np.abs(got[0, 0, 0] - ex[0]).mean()
    print(f"Sod 1024 density L1 vs exact solution: {l1:.3e}", flush=True)
    assert l1 < 2.5e-3


# ---------------------------------------------------------------- Orszag-Tang proxy
def orszag_tang_hydro(n=1024, gamma=5 / 3):
    from gpuimogen_amd.problems import conserved
    L = 2 * np.pi
    x = (np.arange(n) + 0.5) * L / n
    X, Y = np.meshgrid(x, x, indexing="xy")  # [y, x]
    rho = np.full((n, n), gamma ** 2)
    P = np.full((n, n), gamma)
    q = conserved(rho, -np.sin(Y), np.sin(X), np.zeros((n, n)), P, gamma)
    return q.reshape(5, 1, n, n), (L / n, L / n, 1.0)


def test_orszag_tang_hydro_proxy_1024_matches_oracle():
    from gpuimogen_amd import abi, gpuclass as gc, imogen
    n, iters = 1024, 3
    q, h = orszag_tang_hydro(n)
    f = gc.FluidManager(q, 5 / 3, minmass=1e-8, bc_modes=(abi.circular,) * 6)
    loop = imogen.DeviceLoop([f], [abi.METHOD_XINJIN], gc.GeometryManager((n, n, 1), h), cfl=0.8)
    print(f"Orszag-Tang hydro proxy {n}^2 Xin-Jin: {iters} iterations", flush=True)
    loop.run(iters, graph=True)
    t, _ = loop.clock()
    got = f.state
    loop.free()
    f.free()
    ref = OracleFluid(q, 5 / 3, po.XINJIN, (po.BC_CIRCULAR,) * 6)
    tref = oracle_iterations([ref], h, iters, cfl=0.8)
    assert t == pytest.approx(tref, rel=1e-13)
    _check(got, ref.q, 1e-12, "Orszag-Tang hydro proxy vs oracle")
