#!/usr/bin/env python3
"""Benchmark: HLLC fluid step on the 3D Sedov-Taylor blast, FP64, one MI355X.

Workload (the largest single-GPU configuration): 256^3 cells on the
unit cube, rho = 1, v = 0, P = 1e-5, E = 1 deposited in the 123 cells within
r <= 3 dx of the centre, gamma = 5/3, constant-extrapolation boundaries,
HLLC with 2nd-order MUSCL (van Leer limiter), SSP-RK2, CFL 0.4 with the
halved HLLC time step rule. One "step" is one Imogen iteration: the CFL
reduction, then two `fluidstep` calls (forward and backward sweep order),
i.e. 2 cell-updates per cell and 6 directional sweeps.

Prints one JSON line. `value` is cell-updates/s with the state resident in
HBM. `roofline` is the dominant sweep kernel: 80 algorithmic bytes per cell
per directional sweep (5 FP64 conserved variables in, 5 out) divided by its
average launch duration, measured with HIP events on the library's stream
during the timed region. `cpu_baseline` is the FP64 CPU oracle (a restatement
of the reference's arithmetic; the reference has no CPU path) timed on the
host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_SWEEP = 80
METRIC = "HLLC 3D cell-updates/s + % HBM peak; gas-dust speedup learned vs exact drag"
LIMITERS = {"ospre": 0, "vanleer": 1}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(args, q0, h):
    from gpuimogen_amd import abi, gpuclass as gc

    n = q0.shape[-1]
    gamma = 5.0 / 3.0
    lim = LIMITERS[args.limiter]
    f = gc.FluidManager(q0, gamma, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)
    gm = gc.GeometryManager((n, n, n), h)

    state = {"it": 0}

    def iteration():
        tau = gc.cflTimestepFused(f, gm, abi.METHOD_HLLC)
        dt = args.cfl * tau
        it = state["it"]
        gc.fluidstep([f], None, None, None, [dt, 1, 1, it, abi.METHOD_HLLC], gm, limiter=lim)
        gc.fluidstep([f], None, None, None, [dt, 1, -1, it, abi.METHOD_HLLC], gm, limiter=lim)
        state["it"] = it + 1
        return dt

    for _ in range(args.warmup):
        iteration()
    gc.synchronize()
    abi.call("imogen_profile_enable", 1)
    gc.synchronize()
    t0 = time.perf_counter()
    simtime = 0.0
    for k in range(args.steps):
        simtime += 2 * iteration()
        if (k + 1) % max(1, args.steps // 4) == 0:
            log(f"[bench] step {k + 1}/{args.steps}")
    gc.synchronize()
    elapsed = time.perf_counter() - t0
    kern = {}
    names = {0: "hllc_sweep_x", 1: "hllc_sweep_y", 2: "hllc_sweep_z", 3: "boundary", 4: "cfl"}
    for kid, name in names.items():
        cnt = ctypes.c_int64(0)
        tot = ctypes.c_double(0.0)
        abi.call("imogen_profile_read", kid, ctypes.byref(cnt), ctypes.byref(tot))
        kern[name] = {"launches": int(cnt.value), "total_ms": tot.value,
                      "avg_ms": tot.value / cnt.value if cnt.value else None}
    abi.call("imogen_profile_enable", 0)
    final = f.state if args.check else None
    f.free()
    return elapsed, kern, simtime, final


def run_device_loop(args, q0, h):
    """The same iterations with dTime kept on the device and every iteration
    replayed from a HIP graph (imogen.DeviceLoop): no host round trip per CFL
    reduction. Same kernels and bitwise the same results as run_gpu."""
    from gpuimogen_amd import abi, gpuclass as gc, imogen

    n = q0.shape[-1]
    f = gc.FluidManager(q0, 5.0 / 3.0, minmass=1e-8, bc_modes=(abi.extrapConstant,) * 6)
    gm = gc.GeometryManager((n, n, n), h)
    loop = imogen.DeviceLoop([f], [abi.METHOD_HLLC], gm, cfl=args.cfl, limiter=LIMITERS[args.limiter])
    # untimed: at least six iterations, so that all six sweep-order graphs are
    # captured before the timed region whatever --warmup is
    loop.run(max(args.warmup, 6), graph=True)
    gc.synchronize()
    t0 = time.perf_counter()
    done = 0
    chunk = max(1, args.steps // 4)
    while done < args.steps:
        k = min(chunk, args.steps - done)
        loop.run(k, graph=True)
        done += k
        gc.synchronize()
        log(f"[bench] device loop: step {done}/{args.steps}")
    elapsed = time.perf_counter() - t0
    simtime = loop.time
    loop.free()
    f.free()
    return elapsed, simtime


def run_gasdust(args):
    """Gas + dust on a periodic n^3 box (both fluids HLLC; the dust a
    gamma = 1.01 fluid with sound speed 0.3; drag source around each fluid
    step as in imogen.m): the same run with the adaptive drag integrator and
    with the learned surrogate, from the same initial state. Iterations run
    on the device time loop (dTime in device memory, no host round trip per
    iteration), launched eagerly so that the per-kernel events time every
    drag call."""
    from gpuimogen_amd import abi, gpuclass as gc, imogen, learned_drag as L, problems

    n = args.gasdust_n
    gas0, dust0, h, _ = problems.gasdust3d(n, seed=args.seed)
    tg, td = L.cold_h2(), L.iron_dust()
    params, norm, meta = L.load_model()
    if args.drag_resid_tol is not None:
        norm.residTol = args.drag_resid_tol
    L.set_model(params, norm)
    gm = gc.GeometryManager((n, n, n), h)
    out = {}
    for name, method in (("exact", abi.DRAG_EXACT), ("learned", abi.DRAG_LEARNED)):
        fg = gc.FluidManager(gas0, tg.gamma, minmass=1e-8, bc_modes=(abi.circular,) * 6, thermoDetails=tg)
        fd = gc.FluidManager(dust0, td.gamma, minmass=1e-8, bc_modes=(abi.circular,) * 6, thermoDetails=td)
        loop = imogen.DeviceLoop([fg, fd], [abi.METHOD_HLLC, abi.METHOD_HLLC], gm, cfl=0.4, drag_method=method)
        gc.synchronize()
        abi.call("imogen_profile_enable", 1)
        L.learned_stats()
        prof = {"launches": 0, "total_ms": 0.0}

        def take_drag_events():
            # the event pool is read and reset every chunk, so it never runs out
            cnt, tot = ctypes.c_int64(0), ctypes.c_double(0.0)
            abi.call("imogen_profile_read", 5, ctypes.byref(cnt), ctypes.byref(tot))
            prof["launches"] += int(cnt.value)
            prof["total_ms"] += tot.value
            abi.call("imogen_profile_enable", 1)

        t0 = time.perf_counter()
        last = t0
        done = 0
        chunk = 10
        while done < args.gasdust_steps:
            k = min(chunk, args.gasdust_steps - done)
            loop.run(k, graph=False)
            done += k
            take_drag_events()  # synchronises
            now = time.perf_counter()
            if done % max(1, args.gasdust_steps // 4) == 0 or now - last > 20 or done == args.gasdust_steps:
                log(f"[bench] gas-dust {name}: step {done}/{args.gasdust_steps} ({now - t0:.1f} s)")
                last = now
        el = time.perf_counter() - t0
        abi.call("imogen_profile_enable", 0)
        cells, fb = L.learned_stats()
        # six drag calls per iteration: src0001's two half-interval calls in
        # each of the three sources, each source's pair applied by one launch
        drag_it = prof["total_ms"] / args.gasdust_steps
        out[name] = {"seconds": el, "ms_per_step": el / args.gasdust_steps * 1e3,
                     "drag_ms_per_step": drag_it, "drag_ms_per_call": drag_it / 6,
                     "drag_launches_per_step": prof["launches"] / args.gasdust_steps,
                     "sim_time": loop.time, "fallback_share": (fb / cells) if cells else None,
                     "dust_mom": fd.state[2:5].copy()}
        loop.free()
        fg.free()
        fd.free()
    pe, pl = out["exact"].pop("dust_mom"), out["learned"].pop("dust_mom")
    rel_l1 = float(np.abs(pl - pe).sum() / np.abs(pe).sum())
    return {"workload": f"gasdust3d_{n}cubed", "steps": args.gasdust_steps,
            "loop": "device time loop (dTime on the device, drag source around both fluid steps), eager launches",
            "step_speedup": out["exact"]["seconds"] / out["learned"]["seconds"],
            "drag_speedup": out["exact"]["drag_ms_per_call"] / out["learned"]["drag_ms_per_call"],
            "dust_momentum_rel_l1": rel_l1, "exact": out["exact"], "learned": out["learned"],
            "model": {"samples": meta.get("samples"), "epochs": meta.get("epochs"),
                      "residTol": float(norm.residTol)}}


def run_ensemble(args):
    """The batched ensemble: `count` independent 2D four-quadrant HLLC problems
    with log-uniform sizes, per-problem CFL 0.4 on the device, one launch per
    sweep direction and K class over all problems, each iteration a replayed
    hipGraph. First the hydrodynamics alone, then with dust on the odd problems
    (dust-to-gas 0.5) coupled by the learned drag with exact fallback, both
    from the same initial states. The march segment counts are autotuned on the
    ensemble itself. cell_updates_per_s counts the problems' cells (2 per cell
    per iteration); sweep_hbm_frac is the sweeps' algorithmic bytes (80 B per
    cell and sweep, 4 sweeps per iteration, gas and dust) over the whole
    iteration time, against 8 TB/s."""
    from gpuimogen_amd import abi, ensemble as ens, gpuclass as gc, learned_drag as L

    sizes, states = ens.quadrant_problems(args.ensemble_count, seed=args.seed)
    dust = {p: ens.quadrant_dust(states[p], seed=args.seed + 7919 * (p + 1))
            for p in range(len(sizes)) if p % 2 == 1 and not args.ensemble_no_dust}
    e = ens.Ensemble(sizes, sizes, 1.0 / sizes, 1.0 / sizes)
    cells = int(np.sum(sizes.astype(np.int64) ** 2))
    dust_cells = int(sum(int(sizes[p]) ** 2 for p in dust))

    def load(with_dust):
        for p, s in enumerate(states):
            e.set_state(p, s)
        for p, s in (dust.items() if with_dust else ()):
            e.set_dust_state(p, s)

    def timed(label, fluid_cells):
        e.iteration = 0
        e.run(args.warmup, graph=True)
        gc.synchronize()
        t_before = e.times()
        t0 = time.perf_counter()
        done = 0
        chunk = max(1, args.ensemble_steps // 4)
        while done < args.ensemble_steps:
            k = min(chunk, args.ensemble_steps - done)
            e.run(k, graph=True)
            done += k
            log(f"[bench] ensemble ({label}): iteration {done}/{args.ensemble_steps}")
        gc.synchronize()
        el = time.perf_counter() - t0
        adv = e.times() - t_before
        sweep_gbs = fluid_cells * 4 * BYTES_PER_CELL_SWEEP * args.ensemble_steps / el / 1e9
        return el, {"seconds": el, "ms_per_iteration": el / args.ensemble_steps * 1e3,
                    "cell_updates_per_s": cells * 2 * args.ensemble_steps / el,
                    "sweep_bytes_GBs": sweep_gbs, "sweep_hbm_frac": sweep_gbs / HBM_PEAK_GBS,
                    "sim_time_advanced_range": [float(adv.min()), float(adv.max())]}

    # the march segments and X cells per lane are tuned on the hydrodynamics
    # alone (the drag source would only add noise to the timings)
    load(False)
    t0 = time.perf_counter()
    chosen = e.autotune(trials=4)
    xlanes = e.x_lanes()
    tune_s = time.perf_counter() - t0
    load(False)
    _, hydro = timed("hydrodynamics", cells)
    ok = all(bool(np.all(np.isfinite(e.state(p)))) for p in (0, len(sizes) // 2, len(sizes) - 1))
    dusty = None
    if dust:
        params, norm, _ = L.load_model()
        L.set_model(params, norm)
        tg, td = L.cold_h2(), L.iron_dust()
        tg.gamma = 1.4  # the quadrant problems' gas
        e.set_dust([p in dust for p in range(len(sizes))], tg, td, 1.01, abi.DRAG_LEARNED)
        load(True)
        L.learned_stats()
        el, dusty = timed("with dust", cells + dust_cells)
        lc, lf = L.learned_stats()
        ok = ok and all(bool(np.all(np.isfinite(e.dust_state(p)))) for p in list(dust)[:3])
        dusty.update({"problems": len(dust), "dust_to_gas": 0.5, "dust_cells": dust_cells,
                      "drag": "learned + exact fallback",
                      "fluid_cell_updates_per_s": (cells + dust_cells) * 2 * args.ensemble_steps / el,
                      "learned_calls": lc, "fallback_share": lf / lc if lc else None})
    e.free()
    head = dusty if dusty else hydro
    return {"workload": f"ensemble2d_quadrant_{args.ensemble_count}", "problems": len(sizes),
            "sizes": [int(sizes.min()), int(sizes.max())], "cells": cells, "iterations": args.ensemble_steps,
            "warmup": args.warmup, "seconds": head["seconds"], "cell_updates_per_s": head["cell_updates_per_s"],
            "ms_per_iteration": head["ms_per_iteration"],
            "segments_by_ny_class": {"<=128": chosen[0], "<=256": chosen[1], "<=512": chosen[2]},
            "x_cells_per_lane_by_nx_class": {f"<={b}": (f"{k} split" if sp else k) for b, (k, sp)
                                             in zip((64, 128, 192, 256, 368, 496, 512), xlanes) if k},
            "autotune_s": tune_s, "finite": ok, "hydro": hydro, "dust": dusty}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_cpu_baseline(args, q0, h):
    from oracle import pyoracle as po

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    n = q0.shape[-1]
    gamma = 5.0 / 3.0
    lim = po.LIM_VANLEER if args.limiter == "vanleer" else po.LIM_OSPRE

    def sample(iters, nthreads):
        po.set_threads(nthreads)
        q = q0.copy()
        t0 = time.perf_counter()
        for it in range(iters):
            cs = po.soundspeed(q, gamma)
            tau = po.cfl_timestep(q, cs, h, po.HLLC, (n, n, n))
            dt = args.cfl * tau
            for direction in (1, -1):
                po.fluidstep(q, h, dt, direction, it, gamma, po.HLLC, lim, modes=(po.BC_CONST,) * 6)
        return q.shape[1] * q.shape[2] * q.shape[3] * 2 * iters / (time.perf_counter() - t0), time.perf_counter() - t0

    cu, el = sample(args.cpu_iters, threads)
    cu1, el1 = sample(1, 1)
    po.set_threads(threads)
    return {"value": cu, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{args.cpu_iters} iterations (CFL + 2 fluidsteps) of the same {n}^3 Sedov HLLC "
                      f"workload from the same initial state, FP64 oracle, {threads} OpenMP threads, "
                      f"{el:.2f} s",
            "single_thread": {"value": cu1, "unit": "cell-updates/s", "cores": 1,
                              "sample": f"1 iteration of the same workload on one thread, {el1:.2f} s"}}


def run_ospre(args, q0, h):
    """The same Sedov workload with the reference's compiled-in limiter (Ospre,
    gpuclass/cudaFluidStep.cu:132): device-loop throughput and the sweeps'
    HIP-event averages from a short host loop."""
    a = argparse.Namespace(**vars(args))
    a.limiter = "ospre"
    a.steps = max(20, args.steps // 4)
    _, kern, _, _ = run_gpu(a, q0, h)
    el, _ = run_device_loop(a, q0, h)
    cells = q0.shape[-1] ** 3
    sweeps = {k: v["avg_ms"] for k, v in kern.items() if k.startswith("hllc_sweep")}
    dom = max(sweeps, key=sweeps.get)
    return {"limiter": "ospre", "iterations": a.steps, "cell_updates_per_s": cells * 2 * a.steps / el,
            "ms_per_step": el / a.steps * 1e3, "sweep_avg_ms": sweeps,
            "roofline_frac": BYTES_PER_CELL_SWEEP * cells / (sweeps[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS}


def copy_probe(reps=10, ndoubles=1 << 27, quiet=False):
    """Copies 1 GiB `reps` times with the library's copy kernel: the known-size
    reference that calibrates the FETCH_SIZE / WRITE_SIZE counters, and the
    practical HBM ceiling the sweeps are also compared with. Returns GB/s
    (bytes read + written)."""
    from gpuimogen_amd import abi, gpuclass as gc

    a = gc.GPU_Type(dims=(ndoubles, 1, 1))
    b = gc.GPU_Type(dims=(ndoubles, 1, 1))
    abi.call("GPU_copy", b._tp(), a._tp())  # warm
    gc.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        abi.call("GPU_copy", b._tp(), a._tp())
    gc.synchronize()
    el = time.perf_counter() - t0
    nbytes = ndoubles * 8
    gbs = 2 * nbytes * reps / el / 1e9
    if not quiet:
        print(json.dumps({"copy_probe": True, "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                          "launches": reps, "GBs": gbs}), flush=True)
    a.free()
    b.free()
    return gbs


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic_hllc_sweep_sedov256.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--limiter", choices=sorted(LIMITERS), default="vanleer")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="verify the final state is finite and positive")
    ap.add_argument("--copy-probe", action="store_true", help="only run the 1 GiB copy-kernel calibration")
    ap.add_argument("--gasdust-n", type=int, default=192)
    ap.add_argument("--gasdust-steps", type=int, default=1000)
    ap.add_argument("--no-gasdust", action="store_true")
    ap.add_argument("--drag-resid-tol", type=float, default=None,
                    help="override the learned drag's residual tolerance (default: the model file's)")
    ap.add_argument("--ensemble-count", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-no-dust", action="store_true", help="ensemble without the dusty half")
    ap.add_argument("--no-ospre", action="store_true", help="skip the Ospre-limiter Sedov leg")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the library steps "
              "single-GPU canisters (nGPUs == 1) and has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    from gpuimogen_amd import problems

    q0, h, ndep = problems.sedov3d(args.n)
    log(f"[bench] Sedov {args.n}^3, {ndep} deposit cells, {args.steps} timed iterations")
    elapsed_host, kern, simtime, final = run_gpu(args, q0, h)
    cells = args.n ** 3
    cu = cells * 2 * args.steps
    log(f"[bench] Sedov {args.n}^3: device-side time loop, {args.steps} timed iterations")
    elapsed, simtime_dev = run_device_loop(args, q0, h)
    value = cu / elapsed

    sweeps = {k: v for k, v in kern.items() if k.startswith("hllc_sweep") and v["avg_ms"]}
    dom = max(sweeps, key=lambda k: sweeps[k]["total_ms"])
    bytes_per_launch = BYTES_PER_CELL_SWEEP * cells
    achieved = bytes_per_launch / (sweeps[dom]["avg_ms"] * 1e-3) / 1e9
    for k, v in sweeps.items():
        v["achieved_GBs"] = bytes_per_launch / (v["avg_ms"] * 1e-3) / 1e9
    traffic = load_traffic()
    copy_gbs = copy_probe(reps=20, quiet=True)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic.get(dom) if isinstance(traffic, dict) else None,
                "algorithmic_bytes_per_launch": bytes_per_launch,
                "copy_kernel_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs}
    if args.check:
        ok = bool(np.all(np.isfinite(final)) and np.all(final[0] > 0))
        log(f"[bench] final state finite/positive: {ok}")

    ospre = None
    if args.limiter != "ospre" and not args.no_ospre:
        log("[bench] Sedov with the Ospre limiter beside van Leer")
        ospre = run_ospre(args, q0, h)

    gasdust = None
    if not args.no_gasdust:
        log(f"[bench] gas-dust {args.gasdust_n}^3: {args.gasdust_steps} iterations, exact vs learned drag")
        gasdust = run_gasdust(args)

    ensemble = None
    if not args.no_ensemble:
        log(f"[bench] ensemble of {args.ensemble_count} 2D problems: {args.ensemble_steps} iterations")
        ensemble = run_ensemble(args)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU oracle baseline: {args.cpu_iters} iterations")
        cpu = run_cpu_baseline(args, q0, h)

    line = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"sedov3d_hllc_{args.n}cubed", "cells": cells, "method": "HLLC",
                       "limiter": args.limiter, "cfl": args.cfl, "time_scheme": "SSP-RK2",
                       "boundaries": "extrapConstant", "iteration": "CFL + 2 fluidsteps (6 sweeps)",
                       "loop": "device-side dTime, one hipGraph per iteration",
                       "sim_time_advanced": simtime_dev},
            "host_loop": {"cell_updates_per_s": cu / elapsed_host, "ms_per_step": elapsed_host / args.steps * 1e3,
                          "sim_time_advanced": simtime,
                          "note": "CFL read back by the host each iteration; per-kernel times below are from this run"},
            "roofline": roofline, "cpu_baseline": cpu,
            "kernels": kern, "ospre": ospre,
            "drag_speedup_learned_vs_exact": gasdust["step_speedup"] if gasdust else None,
            "gasdust": gasdust, "ensemble": ensemble}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
