"""The standalone C++ driver (gpuimogen_amd/imogen_core, the counterpart of the
reference's imogenCore) runs the same C ABI without Python: its results must
match the oracle run of the same iteration to the fluid-step bar."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpuimogen_amd", "imogen_core")


def run_core(tmp_path, args, ic):
    icf = tmp_path / "ic.bin"
    outf = tmp_path / "out.bin"
    np.ascontiguousarray(ic, dtype=np.float64).tofile(icf)
    cmd = [EXE] + args + ["--ic", str(icf), "--out", str(outf)]
    print("running", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    info = json.loads(res.stdout.strip().split("\n")[-1])
    return np.fromfile(outf, dtype=np.float64).reshape(ic.shape), info


def oracle_run(q, h, steps, gamma, method, limiter, cfl, modes):
    q = q.copy()
    n = q.shape[1:][::-1]
    for it in range(steps):
        cs = po.soundspeed(q, gamma)
        dt = cfl * po.cfl_timestep(q, cs, h, method, n)
        for d in (1, -1):
            po.fluidstep(q, h, dt, d, it, gamma, method, limiter, modes=modes)
    return q


def rel_l1(a, b):
    return [np.abs(a[s] - b[s]).sum() / max(np.abs(b[s]).sum(), 1e-300) for s in range(5)]


def test_core_sedov_matches_oracle(tmp_path):
    from gpuimogen_amd import problems
    q0, h, nd = problems.sedov3d(24)
    got, info = run_core(tmp_path, ["--problem", "sedov", "--n", "24", "--steps", "4", "--method", "hllc",
                                    "--limiter", "1"], q0)
    assert info["deposit_cells"] == nd
    ref = oracle_run(q0, h, 4, 5 / 3, po.HLLC, po.LIM_VANLEER, 0.4, (po.BC_CONST,) * 6)
    err = rel_l1(got, ref)
    print("sedov 24^3, 4 iterations: rel L1", err, flush=True)
    # px/py/pz of a centred blast are antisymmetric: their sums cancel, so the
    # relative measure uses the slab's own magnitude
    assert max(err) <= 1e-12


@pytest.mark.parametrize("method", ["hll", "xinjin"])
def test_core_sod_matches_oracle(tmp_path, method):
    from gpuimogen_amd import problems
    q0, h = problems.sod1d(256)
    om = {"hll": po.HLL, "xinjin": po.XINJIN}[method]
    got, info = run_core(tmp_path, ["--problem", "sod", "--n", "256", "--steps", "20", "--method", method], q0)
    ref = oracle_run(q0, h, 20, 1.4, om, po.LIM_OSPRE, 0.4, (po.BC_CONST,) * 6)
    d = np.abs(got - ref).reshape(5, -1).sum(1)
    scale = np.maximum(np.abs(ref).reshape(5, -1).sum(1), np.abs(ref).sum() * 1e-3)
    print(f"sod {method}: rel L1", d / scale, flush=True)
    assert np.all(d / scale <= 1e-12)


def test_core_gasdust_with_drag_source_matches_oracle(tmp_path):
    """Two fluids through the compiled source step: per iteration the drag
    (LogTrap2) over dt/2, both fluids' fluid step (+1), the drag over dt, (-1),
    the drag over dt/2 (performCompleteTimestep, gpuimogen_core.cu:954-1054,
    with performSourceFunctions type 8), against the oracle's drag and fluid
    step in the same order."""
    from gpuimogen_amd import learned_drag as L, problems
    n, steps = 12, 3
    gas, dust, h, _ = problems.gasdust3d(n, seed=2)
    tg, td = L.cold_h2(), L.iron_dust()
    td.m *= 20.0  # a gentler drag rate keeps 12^3 well inside LogTrap2's accuracy
    thermo = tmp_path / "thermo.bin"
    np.array([getattr(t, f) for t in (tg, td) for f, _ in t._fields_], dtype=np.float64).tofile(thermo)
    q0 = np.stack([gas, dust])
    got, info = run_core(tmp_path, ["--problem", "gasdust", "--n", str(n), "--steps", str(steps), "--drag", "4",
                                    "--thermo", str(thermo)], q0)
    assert info["fluids"] == 2
    ot = [po.Thermo(*[getattr(t, f) for f, _ in t._fields_]) for t in (tg, td)]
    g, d = gas.copy(), dust.copy()
    modes = (po.BC_CIRCULAR,) * 6
    for it in range(steps):
        tau = min(po.cfl_timestep(q, po.soundspeed(q, t.gamma), h, po.HLLC, (n, n, n))
                  for q, t in ((g, tg), (d, td)))
        dt = 0.4 * tau
        po.drag(g, d, ot[0], ot[1], 0.5 * dt, po.DRAG_LOGTRAP2)
        for direction in (1, -1):
            po.fluidstep(g, h, dt, direction, it, tg.gamma, po.HLLC, po.LIM_OSPRE, modes=modes)
            po.fluidstep(d, h, dt, direction, it, td.gamma, po.HLLC, po.LIM_OSPRE, modes=modes)
            po.drag(g, d, ot[0], ot[1], dt if direction > 0 else 0.5 * dt, po.DRAG_LOGTRAP2)
    assert np.all(np.isfinite(got))
    for k, ref in enumerate((g, d)):
        err = rel_l1(got[k], ref)
        print(f"gas+dust fluid {k}: rel L1 {err}", flush=True)
        assert max(err) <= 1e-12
