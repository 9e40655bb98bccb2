#!/usr/bin/env python3
"""Trains the learned-drag network on the GPU and writes the model file.

    python -m gpuimogen_amd.train_drag_model [--samples 1000000] [--epochs 40] [--out PATH]

1. Samples cell states (gas density, dust-to-gas ratio, pressure, relative
   Mach number, step stiffness dt k0) log-uniformly over the ranges of the
   gas-dust and ensemble workloads (learned_drag.sample_cells).
2. Integrates each cell's drag relaxation with the adaptive integrator on the
   GPU to produce features and targets (imogen_drag_training_samples).
3. Trains the 5-64-64-3 network with this library's FP32 forward, loss,
   backward and Adam kernels.
4. Validates on held-out cells: error of ln(|dv(dt)|/|dv(0)|) implied by the
   network, and the share of cells the residual check would accept.
The gas and dust microphysics are cold molecular hydrogen and 10 um iron
grains with the scaling of run/profile_3DXYZ_gasdust.m.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from gpuimogen_amd import learned_drag as L  # noqa: E402
else:
    from . import learned_drag as L


def validate(params, norm, gas, dust, n, seed):
    cells = L.sample_cells(n, gas, np.random.default_rng(seed))
    f, t = L.training_set(cells, gas, dust)
    L.set_model(params, norm)
    xs, _ = L.standardise(f, t, norm)
    y = L.infer_bf16(xs) * np.array(norm.tstd[:]) + np.array(norm.tmean[:])
    tau = np.exp(f[:, 0].astype(np.float64))
    err = np.abs(tau ** 3 * (y[:, 0] - t[:, 0]))  # error in ln r of the mean-rate prediction
    pct = np.percentile(err, [50, 90, 99, 99.9])
    return {"cells": int(len(err)), "lnr_err_p50": pct[0], "lnr_err_p90": pct[1], "lnr_err_p99": pct[2],
            "lnr_err_p999": pct[3], "share_below_tol": float(np.mean(err <= norm.residTol))}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--resid-tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=L.DEFAULT_MODEL)
    args = ap.parse_args()

    gas, dust = L.cold_h2(), L.iron_dust()
    t0 = time.perf_counter()
    cells = L.sample_cells(args.samples, gas, np.random.default_rng(args.seed))
    f, t = L.training_set(cells, gas, dust)
    t_data = time.perf_counter() - t0
    print(f"[train] {len(f)} samples from the adaptive integrator in {t_data:.1f} s", flush=True)
    norm = L.fit_norm(f, t, args.resid_tol)
    t0 = time.perf_counter()
    params, hist = L.train(f, t, norm, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed + 1,
                           log=lambda s: print("[train] " + s, flush=True))
    t_train = time.perf_counter() - t0
    val = validate(params, norm, gas, dust, 100_000, args.seed + 100)
    print("[train] validation: " + json.dumps(val), flush=True)
    meta = {"samples": int(len(f)), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
            "seed": args.seed, "data_seconds": t_data, "train_seconds": t_train, "loss_history": hist,
            "validation": val, "gas": "cold_molecular_hydrogen",
            "dust": "10um_iron_balls, sigma x 0.1, mass x 1e-4"}
    L.save_model(args.out, params, norm, meta)
    print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
