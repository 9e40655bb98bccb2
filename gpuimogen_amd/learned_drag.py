"""Learned gas-dust drag: training set from the exact integrator, training on
the GPU with this library's forward/backward/Adam kernels, model files.

The surrogate replaces the per-cell adaptive integration of the drag
relaxation problem (sourcefunction_2FluidDrag method DRAG_EXACT) inside the
drag source (method DRAG_LEARNED); cells whose residual check fails, or whose
features leave the training box, fall back to the exact integrator on the GPU.
Features, targets and the residual check are described in
include/imogen_abi.h (DragModelNorm) and gpuimogen_amd/csrc/mlp_drag.hip.

Device buffers come from the library itself (imogen_device_alloc and the
copies), so no second GPU runtime is involved; every arithmetic kernel of
training and inference is this library's own. Host-side work is numpy:
sampling cell states, standardisation statistics and epoch permutations.
"""
from __future__ import annotations

import contextlib
import ctypes
import json
import os

import numpy as np

from . import abi

AMU = 1.66053904e-27
KBOLT = 1.381e-23
NPARAMS = abi.IMOGEN_MLP_NPARAMS
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "drag_mlp.json")


def cold_h2():
    """fluidDetailModel('cold_molecular_hydrogen') (astroParams/fluidDetailModel.m)."""
    m = 2.016 * AMU
    mu0 = 8.9135e-6
    sigma = np.sqrt(2 / 27) * 1.7 * np.sqrt(KBOLT * 298.15 * m) / mu0
    return abi.ThermoDetails(5 / 3, m, -1.0, KBOLT, mu0, 0.7, sigma, 0.2)


def iron_dust(sigma_scale=0.1, mass_scale=1e-4):
    """fluidDetailModel('10um_iron_balls') with the cross section and mass scaled
    as run/profile_3DXYZ_gasdust.m:43-46 does."""
    return abi.ThermoDetails(1.01, 3e-11 * mass_scale, 0.1, KBOLT, 0.0, 0.0, np.pi * 25e-12 * sigma_scale, 0.0)


class DeviceArray:
    """A device buffer of the library (imogen_device_alloc), numpy-typed."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p(0)
        abi.call("imogen_device_alloc", self.nbytes, ctypes.byref(p))
        self._p = p.value or 0

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        abi.call("imogen_memcpy_htod", ctypes.c_void_p(d._p), a.ctypes.data_as(ctypes.c_void_p), d.nbytes)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        abi.call("imogen_memcpy_htod", ctypes.c_void_p(self._p), a.ctypes.data_as(ctypes.c_void_p), self.nbytes)

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        abi.call("imogen_memcpy_dtoh", out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self._p), self.nbytes)
        return out

    def ptr(self, offset_elems=0):
        return ctypes.c_void_p(self._p + int(offset_elems) * self.dtype.itemsize)

    def free(self):
        if self._p:
            abi.call("imogen_device_free", ctypes.c_void_p(self._p))
            self._p = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sample_cells(n, gas, rng, tau_range=(1e-3, 0.5), mach_range=(1e-7, 3.0), eps_range=(1e-3, 3.0),
                 rho_range=(0.2, 5.0), p_range=(0.05, 20.0)):
    """Cell states (rho_g, rho_d, |dv|, u_0, dt k0), log-uniform in every
    parameter over ranges that cover the gas-dust and ensemble workloads."""
    def logu(lo, hi, size):
        return np.exp(rng.uniform(np.log(lo), np.log(hi), size))

    rg = logu(*rho_range, n)
    rd = rg * logu(*eps_range, n)
    P = logu(*p_range, n)
    u0 = P / ((gas.gamma - 1) * rg)
    cs = np.sqrt(gas.gamma * (gas.gamma - 1) * u0)
    s0 = cs * logu(*mach_range, n)
    tau = logu(*tau_range, n)
    return np.stack([rg, rd, s0, u0, tau], axis=1)


def training_set(cells, gas, dust):
    """(features, targets) as float32 numpy arrays from the exact integrator on
    the GPU; rows without drag (NaN) are dropped."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    n = cells.shape[0]
    c = DeviceArray.from_host(cells)
    f = DeviceArray((n, 5), np.float32)
    t = DeviceArray((n, 3), np.float32)
    abi.call("imogen_drag_training_samples", ctypes.byref(gas), ctypes.byref(dust), n, c.ptr(), f.ptr(), t.ptr())
    F, T = f.to_host(), t.to_host()
    for d in (c, f, t):
        d.free()
    ok = np.isfinite(F).all(1) & np.isfinite(T).all(1)
    return F[ok], T[ok]


def fit_norm(feat, targ, resid_tol):
    f = np.asarray(feat, dtype=np.float64)
    t = np.asarray(targ, dtype=np.float64)
    norm = abi.DragModelNorm()
    fm, fs, lo, hi = f.mean(0), f.std(0), f.min(0), f.max(0)
    tm, ts = t.mean(0), t.std(0)
    for j in range(5):
        norm.fmean[j], norm.fstd[j], norm.flo[j], norm.fhi[j] = fm[j], fs[j], lo[j], hi[j]
    for j in range(3):
        norm.tmean[j], norm.tstd[j] = tm[j], ts[j]
    norm.residTol = resid_tol
    return norm


def init_params(rng):
    p = np.zeros(NPARAMS, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((5, 64), (64, 64), (64, 3)):
        p[o:o + fan_in * fan_out] = rng.normal(0, np.sqrt(1.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return p


def standardise(feat, targ, norm):
    fm, fs = np.array(norm.fmean[:]), np.array(norm.fstd[:])
    tm, ts = np.array(norm.tmean[:]), np.array(norm.tstd[:])
    return (((feat - fm) / fs).astype(np.float32), ((targ - tm) / ts).astype(np.float32))


def train(feat, targ, norm, epochs=40, batch=2048, lr=3e-3, lr_final=3e-5, seed=0, log=None):
    """Adam on the mean squared error of the standardised targets, with this
    library's FP32 forward / loss / backward / Adam kernels. Returns
    (params, per-epoch mean loss)."""
    rng = np.random.default_rng(seed)
    Xh, Yh = standardise(feat, targ, norm)
    n = Xh.shape[0]
    steps_per_epoch = n // batch
    if steps_per_epoch < 1:
        raise ValueError("fewer samples than one batch")
    X, Y = DeviceArray.from_host(Xh), DeviceArray.from_host(Yh)
    P = DeviceArray.from_host(init_params(rng))
    m = DeviceArray.from_host(np.zeros(NPARAMS, np.float32))
    v = DeviceArray.from_host(np.zeros(NPARAMS, np.float32))
    g = DeviceArray((NPARAMS,), np.float32)
    xb, tb = DeviceArray((batch, 5), np.float32), DeviceArray((batch, 3), np.float32)
    act, y, dy = DeviceArray((batch, 256), np.float32), DeviceArray((batch, 3), np.float32), \
        DeviceArray((batch, 3), np.float32)
    idx = DeviceArray((n,), np.int64)
    loss = DeviceArray((1,), np.float64)
    total = epochs * steps_per_epoch
    history = []
    step = 0
    for ep in range(epochs):
        idx.upload(rng.permutation(n).astype(np.int64))
        loss.upload(np.zeros(1))
        for b in range(steps_per_epoch):
            abi.call("imogen_gather_rows", X.ptr(), idx.ptr(b * batch), batch, 5, xb.ptr())
            abi.call("imogen_gather_rows", Y.ptr(), idx.ptr(b * batch), batch, 3, tb.ptr())
            abi.call("imogen_mlp_forward", P.ptr(), batch, xb.ptr(), act.ptr(), y.ptr())
            abi.call("imogen_mlp_mse_grad", batch, y.ptr(), tb.ptr(), dy.ptr(), loss.ptr())
            abi.call("imogen_mlp_backward", P.ptr(), batch, xb.ptr(), act.ptr(), dy.ptr(), g.ptr())
            step += 1
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * step / total))
            abi.call("imogen_mlp_adam", P.ptr(), g.ptr(), m.ptr(), v.ptr(), NPARAMS, lr_t, 0.9, 0.999, 1e-8, step)
        history.append(float(loss.to_host()[0]) / steps_per_epoch)
        if log:
            log(f"epoch {ep + 1}/{epochs}: mse {history[-1]:.3e}")
    params = P.to_host()
    for d in (X, Y, P, m, v, g, xb, tb, act, y, dy, idx, loss):
        d.free()
    return params, history


def infer_bf16(feat_std):
    """Standardised outputs of the loaded model through the BF16 MFMA kernel."""
    x = DeviceArray.from_host(np.asarray(feat_std, dtype=np.float32))
    y = DeviceArray((x.shape[0], 3), np.float32)
    abi.call("imogen_mlp_infer_bf16", x.shape[0], x.ptr(), y.ptr())
    out = y.to_host()
    x.free()
    y.free()
    return out


def set_model(params, norm):
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.size != NPARAMS:
        raise ValueError(f"expected {NPARAMS} parameters, got {p.size}")
    abi.call("imogen_drag_set_model", p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(norm))


def norm_to_dict(norm):
    return {f: [float(x) for x in getattr(norm, f)] if f != "residTol" else float(norm.residTol)
            for f, _ in abi.DragModelNorm._fields_}


def norm_from_dict(d):
    norm = abi.DragModelNorm()
    for f, _ in abi.DragModelNorm._fields_:
        if f == "residTol":
            norm.residTol = d[f]
        else:
            arr = getattr(norm, f)
            for j, x in enumerate(d[f]):
                arr[j] = x
    return norm


def save_model(path, params, norm, meta=None):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        json.dump({"params": [float(x) for x in np.asarray(params, dtype=np.float32)], "norm": norm_to_dict(norm),
                   "meta": meta or {}}, fh)


def load_model(path=DEFAULT_MODEL):
    with open(path) as fh:
        d = json.load(fh)
    return np.asarray(d["params"], dtype=np.float32), norm_from_dict(d["norm"]), d.get("meta", {})


def learned_stats():
    cells = ctypes.c_int64(0)
    fb = ctypes.c_int64(0)
    abi.call("imogen_drag_learned_stats", ctypes.byref(cells), ctypes.byref(fb))
    return cells.value, fb.value
