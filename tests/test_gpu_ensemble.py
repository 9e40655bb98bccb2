"""The batched ensemble path (imogen_ensemble_*) against the single-array path
of the same library, which is itself parity-tested against the oracle: every
problem, advanced inside the ensemble with its own device-side CFL step, must
match the same problem advanced alone, and graph replay must match eager
launches bitwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [64, 77, 100, 128, 150, 200, 256, 300]


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
    # after tuning (other segment counts) the ensemble still matches the single path
    for p, s in enumerate(states):
        a.set_state(p, s)
    a.iteration = 0
    a.run(3, graph=True)
    ref, _ = single_path(states[3], 3, SIZES[3])
    assert np.abs(a.state(3) - ref).max() <= 1e-12 * np.abs(ref).max()
    a.free()
    b.free()
