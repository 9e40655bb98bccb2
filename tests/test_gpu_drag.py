"""GPU parity of the drag source and the learned-drag network kernels.

Bars: LogTrap2/3 match the oracle to 1e-12 relative L1 per slab (the same
arithmetic, FMA contraction aside). The adaptive integrator matches to 1e-10:
its step-size controller sees FP64 rounding differences, so both sides stop
within the integration tolerance (1e-12 on ln|dv|) of the true solution but
not necessarily on the same step sequence. The FP32 training kernels match the
FP64 numpy restatement (oracle/mlp_oracle.py) to FP32 rounding; BF16 inference
matches a BF16-rounding emulation of the same network.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from oracle import pyoracle as po

from .test_gpu_hllc import gc
from .test_oracle_drag import gas_dust_thermo

pytestmark = pytest.mark.gpu


def abi():
    from gpuimogen_amd import abi as a
    return a


def ld():
    from gpuimogen_amd import learned_drag
    return learned_drag


def thermo_pair():
    a = abi()
    tg, td = gas_dust_thermo()
    return (a.ThermoDetails(*[getattr(tg, f) for f, _ in tg._fields_]),
            a.ThermoDetails(*[getattr(td, f) for f, _ in td._fields_]))


def random_pair(sh, seed, eps_median=0.1):
    rng = np.random.default_rng(seed)
    g = np.zeros((5,) + sh)
    d = np.zeros((5,) + sh)
    g[0] = rng.uniform(0.5, 2, sh)
    d[0] = g[0] * np.exp(rng.normal(np.log(eps_median), 1.0, sh))
    vg = rng.normal(0, 0.4, (3,) + sh)
    vd = rng.normal(0, 0.4, (3,) + sh)
    vd[:, 0, 0, :3] = vg[:, 0, 0, :3]  # no relative motion
    g[2:] = g[0] * vg
    d[2:] = d[0] * vd
    g[1] = rng.uniform(0.5, 2, sh) / (5 / 3 - 1) + .5 * g[0] * (vg ** 2).sum(0)
    d[1] = .5 * d[0] * (vd ** 2).sum(0)
    return g, d


def compare_slabs(got, ref, tol, what):
    for s in range(5):
        dsum = np.abs(got[s] - ref[s]).sum()
        scale = np.abs(ref[s]).sum()
        rel = dsum / scale if scale > 0 else dsum
        print(f"{what} slab {s}: rel L1 {rel:.3e}", flush=True)
        assert rel <= tol, (what, s, rel)


def run_gpu_drag(g, d, dt, method, sh):
    a = abi()
    tg, td = thermo_pair()
    fg = gc().FluidManager(g, tg.gamma, thermoDetails=tg)
    fd = gc().FluidManager(d, td.gamma, thermoDetails=td)
    nz, ny, nx = sh
    geo = gc().GeometryManager((nx, ny, nz), (0.1, 0.1, 0.1))
    gc().cudaSource2FluidDrag([fg, fd], geo, [dt, method])
    out = fg.state, fd.state
    fg.free()
    fd.free()
    return out


@pytest.mark.parametrize("method,tol", [(4, 1e-12), (5, 1e-12), (6, 1e-10)])
def test_drag_parity(method, tol):
    sh = (6, 10, 70)
    g, d = random_pair(sh, seed=method)
    print(f"drag method {method} shape {sh}", flush=True)
    gg, dd = run_gpu_drag(g, d, 0.02, method, sh)
    tg, td = gas_dust_thermo()
    rg, rd = g.copy(), d.copy()
    po.drag(rg, rd, tg, td, 0.02, method)
    compare_slabs(gg, rg, tol, f"gas m{method}")
    compare_slabs(dd, rd, tol, f"dust m{method}")


@pytest.mark.parametrize("method", [0, 1, 2])
def test_drag_emp_rk4_parity(method):
    """Explicit midpoint / RK4 / exponential Euler (solveDragEMP, solveDragRK4,
    solveDragETDRK1). Like the reference they give NaN at |dv| = 0 (and EMP /
    RK4 when a stage |dv| turns negative), so the state has relative motion
    everywhere and dt keeps k dt <= 1/2."""
    sh = (5, 9, 66)
    g, d = random_pair(sh, seed=20 + method)
    d[2:, 0, 0, :3] = d[2:, 0, 0, :3] * 0.5 + 0.01
    tg, td = gas_dust_thermo()
    P9 = po.drag_params(tg, td)
    dv = g[2:] / g[0] - d[2:] / d[0]
    s0 = np.sqrt((dv ** 2).sum(0))
    u0 = (g[1] - .5 * (g[2:] ** 2).sum(0) / g[0]) / g[0]
    kmax = max(po.drag_k(P9, a, b, c, e) for a, b, c, e in zip(g[0].ravel(), d[0].ravel(), s0.ravel(), u0.ravel()))
    dt = 0.5 / kmax
    print(f"drag method {method} shape {sh} dt {dt:.3e}", flush=True)
    gg, dd = run_gpu_drag(g, d, dt, method, sh)
    rg, rd = g.copy(), d.copy()
    po.drag(rg, rd, tg, td, dt, method)
    assert np.all(np.isfinite(rg)) and np.all(np.isfinite(rd))
    compare_slabs(gg, rg, 1e-12, f"gas m{method}")
    compare_slabs(dd, rd, 1e-12, f"dust m{method}")


def test_unbuilt_drag_methods_fail_loudly():
    sh = (2, 3, 4)
    g, d = random_pair(sh, seed=1)
    for m in (3,):
        with pytest.raises(abi().ImogenError):
            run_gpu_drag(g, d, 0.01, m, sh)


@pytest.mark.parametrize("method", [4, 5])
def test_source_step_two_fluids_is_one_full_dt_drag_call(method):
    """performSourceFunctions type 8 without radiation (src___D,
    sourceStep.cu:300-326) is one sourcefunction_2FluidDrag call over
    dTime * tFraction: bitwise the direct call, and the oracle to 1e-12."""
    sh = (4, 6, 40)
    g, d = random_pair(sh, seed=30 + method)
    nz, ny, nx = sh
    dTime, frac = 0.04, 0.5
    print(f"sourceStep drag method {method} shape {sh}", flush=True)
    a = abi()
    tg, td = thermo_pair()
    geo = gc().GeometryManager((nx, ny, nz), (0.1, 0.1, 0.1))
    fg = gc().FluidManager(g, tg.gamma, thermoDetails=tg)
    fd = gc().FluidManager(d, td.gamma, thermoDetails=td)
    gc().sourceStep([fg, fd], geo, dTime, [1, 0, a.METHOD_HLLC, frac], multifluidDragMethod=method)
    got_g, got_d = fg.state, fd.state
    fg.free()
    fd.free()
    ref_g, ref_d = run_gpu_drag(g, d, dTime * frac, method, sh)
    assert np.all(np.isfinite(got_g)) and np.all(np.isfinite(got_d))
    assert np.array_equal(got_g, ref_g) and np.array_equal(got_d, ref_d)
    og, od = gas_dust_thermo()
    rg, rd = g.copy(), d.copy()
    po.drag(rg, rd, og, od, dTime * frac, method)
    compare_slabs(got_g, rg, 1e-12, f"gas m{method}")
    compare_slabs(got_d, rd, 1e-12, f"dust m{method}")


def test_source_step_one_fluid_and_unbuilt_terms():
    a = abi()
    sh = (2, 3, 8)
    g, _ = random_pair(sh, seed=2)
    tg, _ = thermo_pair()
    geo = gc().GeometryManager((8, 3, 2), (0.1, 0.1, 0.1))
    fg = gc().FluidManager(g, tg.gamma, thermoDetails=tg)
    gc().sourceStep([fg], geo, 0.01, [1, 0, a.METHOD_HLLC, 1.0])  # type 0: nothing to do
    assert np.array_equal(fg.state, g)
    with pytest.raises(a.ImogenError) as e:
        gc().sourceStep([fg], geo, 0.01, [1, 0, a.METHOD_HLLC, 1.0], radiation=(0.5, 1.0))
    assert e.value.code == a.ERROR_NOIMPLEMENT
    assert np.array_equal(fg.state, g)
    fg.free()


# ---------------------------------------------------------------- network kernels
def _dev(x, dtype=np.float32):
    return ld().DeviceArray.from_host(np.asarray(x, dtype=dtype))


def test_mlp_forward_backward_match_oracle():
    L = ld()
    rng = np.random.default_rng(3)
    p = L.init_params(rng)
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 64)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, 64)
    p[mo.OFF_B3:] = rng.normal(0, 0.1, 3)
    B = 1000  # not a multiple of the 64-row block
    x = rng.normal(0, 1, (B, 5)).astype(np.float32)
    dy = rng.normal(0, 1, (B, 3)).astype(np.float32)
    P, X, DY = _dev(p), _dev(x), _dev(dy)
    act = L.DeviceArray((B, 256), np.float32)
    y = L.DeviceArray((B, 3), np.float32)
    gr = L.DeviceArray((mo.NPARAMS,), np.float32)
    abi().call("imogen_mlp_forward", P.ptr(), B, X.ptr(), act.ptr(), y.ptr())
    abi().call("imogen_mlp_backward", P.ptr(), B, X.ptr(), act.ptr(), DY.ptr(), gr.ptr())
    y, act, gr = y.to_host(), act.to_host(), gr.to_host()
    yr, (z1, h1, z2, h2) = mo.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.abs(y - yr).max() <= 1e-5 * np.abs(yr).max()
    for k, ref in enumerate((z1, h1, z2, h2)):
        assert np.abs(act[:, 64 * k:64 * (k + 1)] - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    gref = mo.backward(p.astype(np.float64), x.astype(np.float64), (z1, h1, z2, h2), dy.astype(np.float64))
    for name, lo, hi in (("W1", mo.OFF_W1, mo.OFF_B1), ("b1", mo.OFF_B1, mo.OFF_W2), ("W2", mo.OFF_W2, mo.OFF_B2),
                         ("b2", mo.OFF_B2, mo.OFF_W3), ("W3", mo.OFF_W3, mo.OFF_B3), ("b3", mo.OFF_B3, mo.NPARAMS)):
        err = np.abs(gr[lo:hi] - gref[lo:hi]).max() / np.abs(gref[lo:hi]).max()
        print(f"grad {name}: max rel {err:.2e}", flush=True)
        assert err <= 2e-5, name


def test_loss_gradient_gather_and_adam_match_oracle():
    L = ld()
    rng = np.random.default_rng(4)
    # mean-squared-error gradient and loss
    B = 777
    y, t = rng.normal(size=(B, 3)).astype(np.float32), rng.normal(size=(B, 3)).astype(np.float32)
    Y, T = _dev(y), _dev(t)
    DY = L.DeviceArray((B, 3), np.float32)
    loss = _dev(np.zeros(1), np.float64)
    abi().call("imogen_mlp_mse_grad", B, Y.ptr(), T.ptr(), DY.ptr(), loss.ptr())
    e = y.astype(np.float64) - t
    assert np.allclose(DY.to_host(), 2 * e / e.size, rtol=1e-6, atol=1e-9)
    # the kernel squares the FP32 difference
    assert loss.to_host()[0] == pytest.approx((e ** 2).mean(), rel=1e-6)
    # row gather
    src = rng.normal(size=(50, 5)).astype(np.float32)
    idx = rng.integers(0, 50, 31).astype(np.int64)
    S, I = _dev(src), _dev(idx, np.int64)
    D = L.DeviceArray((31, 5), np.float32)
    abi().call("imogen_gather_rows", S.ptr(), I.ptr(), 31, 5, D.ptr())
    assert np.array_equal(D.to_host(), src[idx])
    # Adam
    n = abi().IMOGEN_MLP_NPARAMS
    p, g = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    m, v = rng.normal(size=n).astype(np.float32) * 0.1, rng.uniform(0, 0.1, n).astype(np.float32)
    P, G, M, V = _dev(p), _dev(g), _dev(m), _dev(v)
    abi().call("imogen_mlp_adam", P.ptr(), G.ptr(), M.ptr(), V.ptr(), n, 1e-2, 0.9, 0.999, 1e-8, 7)
    got = P.to_host(), M.to_host(), V.to_host()
    ref = mo.adam(p.astype(np.float64), g.astype(np.float64), m.astype(np.float64), v.astype(np.float64),
                  1e-2, 0.9, 0.999, 1e-8, 7)
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6)


def test_bf16_inference_matches_emulation():
    L = ld()
    rng = np.random.default_rng(5)
    p = L.init_params(rng)
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 64)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, 64)
    p[mo.OFF_B3:] = rng.normal(0, 0.1, 3)
    norm = abi().DragModelNorm()
    for j in range(5):
        norm.fstd[j] = 1.0
    L.set_model(p, norm)
    B = 1000
    x = rng.normal(0, 1, (B, 5)).astype(np.float32)
    y = L.infer_bf16(x)
    yb = mo.forward_bf16(p, x)
    yf, _ = mo.forward(p.astype(np.float64), x.astype(np.float64))
    scale = np.abs(yf).max()
    err_emul = np.abs(y - yb).max() / scale
    err_fp = np.abs(y - yf).max() / scale
    print(f"bf16 inference: vs emulation {err_emul:.2e}, vs fp64 network {err_fp:.2e}", flush=True)
    assert err_emul <= 1e-2
    assert err_fp <= 5e-2


# ---------------------------------------------------------------- learned drag
def test_learned_drag_fallback_paths():
    """residTol = 0 sends every cell to the exact integrator (result identical
    to method 6); an unbounded tolerance and box accept every moving cell."""
    L = ld()
    rng = np.random.default_rng(6)
    p = L.init_params(rng) * 0.0  # predicts tmean: a crude but valid model
    norm = abi().DragModelNorm()
    for j in range(5):
        norm.fstd[j] = 1.0
        norm.flo[j] = -1e30
        norm.fhi[j] = 1e30
    sh = (4, 6, 50)
    g, d = random_pair(sh, seed=8)
    n_moving = int((np.sqrt(((g[2:] / g[0] - d[2:] / d[0]) ** 2).sum(0)) >= 1e-9).sum())
    norm.residTol = 0.0
    L.set_model(p, norm)
    L.learned_stats()
    gl, dl = run_gpu_drag(g, d, 0.02, 7, sh)
    cells, fb = L.learned_stats()
    assert cells == g[0].size and fb == n_moving
    ge, de = run_gpu_drag(g, d, 0.02, 6, sh)
    # same integrator, compiled into two kernels (FMA contraction may differ)
    assert np.allclose(gl, ge, rtol=1e-13, atol=0) and np.allclose(dl, de, rtol=1e-13, atol=0)
    # with dt k0 <= 0.1 the baseline rate stays positive and every moving cell
    # is accepted under an unbounded tolerance
    norm.residTol = 1e30
    L.set_model(p, norm)
    run_gpu_drag(g, d, 1e-3, 7, sh)
    cells, fb = L.learned_stats()
    assert fb == 0
    # A zero remainder leaves k/k0 = 1 + tau c g: with dt = 1e-7 (tau = dt k0
    # <= 1e-4 here) every cell passes a 1e-10 residual check and matches the
    # exact integrator to O(tau^3); a wrong initial slope g would miss by
    # O(tau^2).
    norm.residTol = 1e-10
    L.set_model(p, norm)
    gl, dl = run_gpu_drag(g, d, 1e-7, 7, sh)
    cells, fb = L.learned_stats()
    assert fb == 0
    ge, de = run_gpu_drag(g, d, 1e-7, 6, sh)
    dvl = gl[2:] / gl[0] - dl[2:] / dl[0]
    dve = ge[2:] / ge[0] - de[2:] / de[0]
    assert np.abs(dvl - dve).max() <= 1e-10 * np.abs(dve).max()


def cells_to_grid(cells):
    """Gas/dust states on a 1 x 1 x n grid from sampled (rho_g, rho_d, |dv|,
    u_0, .) rows: the gas moves along x, the dust is at rest."""
    n = cells.shape[0]
    g = np.zeros((5, 1, 1, n))
    d = np.zeros((5, 1, 1, n))
    rg, rd, s0, u0 = cells[:, 0], cells[:, 1], cells[:, 2], cells[:, 3]
    g[0, 0, 0] = rg
    g[2, 0, 0] = rg * s0
    g[1, 0, 0] = u0 * rg + .5 * rg * s0 * s0
    d[0, 0, 0] = rd
    return g, d


def test_training_on_gpu_reduces_loss():
    """A short training run with this library's kernels: data from the
    adaptive integrator, FP32 forward / loss / backward / Adam on the GPU."""
    L = ld()
    tg, td = thermo_pair()
    cells = L.sample_cells(200_000, tg, np.random.default_rng(11))
    f, t = L.training_set(cells, tg, td)
    assert len(f) > 190_000
    norm = L.fit_norm(f, t, resid_tol=1e-8)
    params, hist = L.train(f, t, norm, epochs=12, batch=2048, lr=3e-3, seed=1, log=print)
    assert np.all(np.isfinite(params))
    assert hist[-1] < 0.3 * hist[0], hist


def test_trained_model_cuts_fallbacks_and_stays_accurate():
    """The committed network (gpuimogen_amd/models/drag_mlp.json, trained on
    the GPU by gpuimogen_amd/train_drag_model.py) against the same learned
    drag with an untrained (zero) network, on held-out cells whose step
    stiffness dt k0 lies in [1e-3, 0.3]: the trained network passes the
    residual check on far more cells, and every cell, accepted or not, stays
    within 1e-6 of the adaptive integrator in ln|dv|."""
    L = ld()
    tg, td = thermo_pair()
    params, norm, meta = L.load_model()
    held = L.sample_cells(60_000, tg, np.random.default_rng(12))
    P9 = po.drag_params(*gas_dust_thermo())
    dt = 2e-3
    k0 = np.array([po.drag_k(P9, r[0], r[1], r[2], r[3]) for r in held])
    keep = (k0 * dt >= 1e-3) & (k0 * dt <= 0.3)
    held = held[keep][:20_000]
    g, d = cells_to_grid(held)
    sh = g.shape[1:]
    ge, de = run_gpu_drag(g, d, dt, 6, sh)
    dv_e = ge[2] / ge[0] - de[2] / de[0]
    counts = {}
    for name, p in (("zero", np.zeros_like(params)), ("trained", params)):
        L.set_model(p, norm)
        L.learned_stats()
        gl, dl = run_gpu_drag(g, d, dt, 7, sh)
        ncell, fb = L.learned_stats()
        dv_l = gl[2] / gl[0] - dl[2] / dl[0]
        err = np.abs(np.log(dv_l / dv_e))
        counts[name] = fb
        print(f"{name} network: {fb} of {ncell} cells fell back; max |ln r error| {err.max():.2e}", flush=True)
        assert err.max() <= 1e-6
    assert counts["trained"] < 0.5 * counts["zero"], counts


# ---------------------------------------------------------------- invalid states
def _poisoned_pair(seed=40):
    """4 x 6 x 50 gas/dust pair with three cells the drag is not defined on:
    a NaN gas cell, a moving cell with zero gas internal energy, and a cell
    with negative dust density."""
    sh = (4, 6, 50)
    g, d = random_pair(sh, seed=seed)
    bad_g, bad_d = g.copy(), d.copy()
    bad_g[:, 1, 2, 3] = np.nan
    bad_g[1, 2, 4, 10] = 0.5 * (bad_g[2:, 2, 4, 10] ** 2).sum() / bad_g[0, 2, 4, 10]
    bad_d[0, 3, 5, 49] = -0.5
    mask = np.ones(sh, bool)
    mask[1, 2, 3] = mask[2, 4, 10] = mask[3, 5, 49] = False
    return sh, g, d, bad_g, bad_d, mask


@pytest.mark.parametrize("method", [4, 6, 7])
def test_invalid_cells_fail_promptly_and_leave_the_rest_exact(method):
    """Regression: the adaptive integrator used to spin to its 4096-step
    budget on NaN cells (a NaN error estimate grew the step forever) and the
    overrun was dropped. Now every drag path counts such cells into the device
    error word, sourcefunction_2FluidDrag returns ERROR_CRASH at once, and the
    valid cells get exactly what they get without the invalid ones."""
    import time
    a = abi()
    sh, g, d, bad_g, bad_d, mask = _poisoned_pair()
    if method == 7:
        L = ld()
        params, norm, _ = L.load_model()
        L.set_model(params, norm)
    print(f"invalid drag cells, method {method}, shape {sh}", flush=True)
    ok_g, ok_d = run_gpu_drag(g, d, 0.02, method, sh)
    tg, td = thermo_pair()
    fg = gc().FluidManager(bad_g, tg.gamma, thermoDetails=tg)
    fd = gc().FluidManager(bad_d, td.gamma, thermoDetails=td)
    geo = gc().GeometryManager((sh[2], sh[1], sh[0]), (0.1, 0.1, 0.1))
    t0 = time.perf_counter()
    with pytest.raises(a.ImogenError) as e:
        gc().cudaSource2FluidDrag([fg, fd], geo, [0.02, method])
    took = time.perf_counter() - t0
    assert e.value.code == a.ERROR_CRASH
    assert took < 2.0, took
    got_g, got_d = fg.state, fd.state
    fg.free()
    fd.free()
    assert np.array_equal(got_g[:, mask], ok_g[:, mask]) and np.array_equal(got_d[:, mask], ok_d[:, mask])
    assert a.lib().imogen_check_errors() == 0  # the report cleared the record
    # the oracle flags the same state (exact integrator)
    if method == 6:
        og, od = gas_dust_thermo()
        rg, rd = bad_g.copy(), bad_d.copy()
        with pytest.raises(po.DragStateError):
            po.drag(rg, rd, og, od, 0.02, 6)


def test_device_loop_reports_invalid_state_and_stops_the_clock():
    """A gas-dust device loop on a state with invalid cells returns
    ERROR_CRASH from imogen_timeloop_run; an all-NaN fluid has no finite CFL
    limit, so its clock takes no time step."""
    from gpuimogen_amd import imogen as im
    a = abi()
    sh, g, d, bad_g, bad_d, _ = _poisoned_pair(41)
    tg, td = thermo_pair()
    fg = gc().FluidManager(bad_g, tg.gamma, thermoDetails=tg, bc_modes=(a.circular,) * 6)
    fd = gc().FluidManager(bad_d, td.gamma, thermoDetails=td, bc_modes=(a.circular,) * 6)
    geo = gc().GeometryManager((sh[2], sh[1], sh[0]), (0.1, 0.1, 0.1))
    loop = im.DeviceLoop([fg, fd], [a.METHOD_HLLC, a.METHOD_HLLC], geo, drag_method=a.DRAG_LOGTRAP2)
    print("device loop on an invalid gas-dust state", flush=True)
    with pytest.raises(a.ImogenError) as e:
        loop.run(3, graph=False)
    assert e.value.code == a.ERROR_CRASH
    loop.free()
    fg.free()
    fd.free()
    nan = np.full((5,) + sh, np.nan)
    f = gc().FluidManager(nan, 5 / 3, bc_modes=(a.circular,) * 6)
    loop = im.DeviceLoop([f], [a.METHOD_HLLC], geo)
    print("device loop on an all-NaN state", flush=True)
    with pytest.raises(a.ImogenError) as e:
        loop.run(2, graph=True)
    assert e.value.code == a.ERROR_CRASH
    t, dt = loop.clock()
    assert t == 0.0 and dt == 0.0
    loop.free()
    f.free()
