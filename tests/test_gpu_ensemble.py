"""The batched ensemble path (imogen_ensemble_*) against the single-array path
of the same library, which is itself parity-tested against the oracle: every
problem, advanced inside the ensemble with its own device-side CFL step, must
match the same problem advanced alone, and graph replay must match eager
launches bitwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# every X line class: whole lines of 1-8 cells per lane, and lines split into
# halves (257-368 cells: 3 per lane; 385-496: 4 per lane), with both edges of
# the split classes and odd lengths
SIZES = [64, 77, 100, 128, 150, 200, 256, 300, 368, 369, 385, 451, 496, 509]


def make(seed=3):
    from gpuimogen_amd import ensemble as ens
    rng = np.random.default_rng(seed)
    states = []
    for n in SIZES:
        _, st = ens.quadrant_problems(1, seed=int(rng.integers(1 << 30)), nmin=n, nmax=n)
        states.append(st[0])
    return states


def single_path(state, iterations, n):
    from gpuimogen_amd import abi, gpuclass as gc
    f = gc.FluidManager(state.reshape(5, 1, n, n), 1.4, bc_modes=(abi.extrapConstant,) * 6)
    gm = gc.GeometryManager((n, n, 1), (1.0 / n, 1.0 / n, 1.0))
    t = 0.0
    for it in range(iterations):
        dt = 0.4 * gc.cflTimestepFused(f, gm, abi.METHOD_HLLC)
        gc.fluidstep([f], None, None, None, [dt, 1, 1, it, abi.METHOD_HLLC], gm)
        gc.fluidstep([f], None, None, None, [dt, 1, -1, it, abi.METHOD_HLLC], gm)
        t += 2 * dt
    out = f.state.reshape(5, n, n)
    f.free()
    return out, t


def build(states):
    from gpuimogen_amd import ensemble as ens
    n = np.array(SIZES)
    e = ens.Ensemble(n, n, 1.0 / n, 1.0 / n)
    for p, s in enumerate(states):
        e.set_state(p, s)
    return e


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_matches_single_problems(graph):
    states = make()
    iters = 7  # more than one sweep-order period of graphs
    e = build(states)
    e.run(iters, graph=graph)
    times = e.times()
    for p, n in enumerate(SIZES):
        print(f"problem {p}: {n}^2, graph={graph}", flush=True)
        ref, tref = single_path(states[p], iters, n)
        got = e.state(p)
        for s in range(5):
            d = np.abs(got[s] - ref[s]).sum()
            scale = max(np.abs(ref[s]).sum(), 1e-300)
            assert d / scale <= 1e-12, (p, s, d / scale)
        assert times[p] == pytest.approx(tref, rel=1e-13)
    e.free()


def test_graph_replay_equals_eager_and_autotune_runs():
    states = make(seed=5)
    a, b = build(states), build(states)
    a.run(8, graph=False)
    b.run(8, graph=True)
    for p in range(len(SIZES)):
        assert np.array_equal(a.state(p), b.state(p))
    chosen = a.autotune(trials=1)
    assert all(c in (1, 2, 4) for c in chosen)
    # cells per lane per nx group: the least that fits, or up to two more
    lanes = a.x_lanes()
    least = {0: 1, 1: 2, 2: 3, 3: 4, 4: 3, 5: 4, 6: 5}
    for g, (k, sp) in enumerate(lanes):
        assert k in range(least[g], least[g] + 3) and sp == (g >= 4), (g, k, sp)
    # after tuning (other segment counts and cells per lane) every problem still
    # matches the single path
    for p, s in enumerate(states):
        a.set_state(p, s)
    a.iteration = 0
    a.run(3, graph=True)
    for p, n in enumerate(SIZES):
        print(f"problem {p}: {n}^2 after tuning {chosen} {lanes}", flush=True)
        ref, _ = single_path(states[p], 3, n)
        assert np.abs(a.state(p) - ref).max() <= 1e-12 * np.abs(ref).max(), p
    a.free()
    b.free()


DUSTY_SIZES = [64, 77, 100, 130]


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_with_dust_matches_single_problems(graph):
    """Odd problems carry dust (dust-to-gas ratio 0.5, learned drag with exact
    fallback); every problem must match the same problem advanced alone by the
    host-driven loop with the same fluids and drag source (imogen.Simulation)."""
    from gpuimogen_amd import abi, ensemble as ens, gpuclass as gc, imogen, learned_drag as L
    params, norm, _ = L.load_model()
    L.set_model(params, norm)
    tg, td = L.cold_h2(), L.iron_dust()
    tg.gamma = 1.4  # the quadrant problems' gas
    rng = np.random.default_rng(21)
    gas, dust = [], {}
    for p, n in enumerate(DUSTY_SIZES):
        _, st = ens.quadrant_problems(1, seed=int(rng.integers(1 << 30)), nmin=n, nmax=n)
        gas.append(st[0])
        if p % 2:
            dust[p] = ens.quadrant_dust(st[0], seed=100 + p)
    sizes = np.array(DUSTY_SIZES)
    e = ens.Ensemble(sizes, sizes, 1.0 / sizes, 1.0 / sizes)
    for p, s in enumerate(gas):
        e.set_state(p, s)
    e.set_dust([p % 2 == 1 for p in range(len(sizes))], tg, td, 1.01, abi.DRAG_LEARNED)
    for p, s in dust.items():
        e.set_dust_state(p, s)
    iters = 3
    e.run(iters, graph=graph)
    times = e.times()
    for p, n in enumerate(DUSTY_SIZES):
        print(f"problem {p}: {n}^2, dust={p in dust}, graph={graph}", flush=True)
        gm = gc.GeometryManager((n, n, 1), (1.0 / n, 1.0 / n, 1.0))
        fl = [gc.FluidManager(gas[p].reshape(5, 1, n, n), 1.4, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6,
                              thermoDetails=tg)]
        if p in dust:
            fl.append(gc.FluidManager(dust[p].reshape(5, 1, n, n), 1.01, minmass=1e-8,
                                      bc_modes=(abi.extrapConstant,) * 6, thermoDetails=td))
        sim = imogen.Simulation(fl, [abi.METHOD_HLLC] * len(fl), gm, cfl=0.4,
                                drag_method=abi.DRAG_LEARNED if p in dust else None)
        for _ in range(iters):
            sim.step()
        pairs = [(e.state(p), fl[0].state.reshape(5, n, n))]
        if p in dust:
            pairs.append((e.dust_state(p), fl[1].state.reshape(5, n, n)))
        for got, ref in pairs:
            for s in range(5):
                d = np.abs(got[s] - ref[s]).sum()
                scale = max(np.abs(ref[s]).sum(), 1e-300)
                assert d / scale <= 1e-12, (p, s, d / scale)
        assert times[p] == pytest.approx(sim.time, rel=1e-13)
        for f in fl:
            f.free()
    e.free()
