"""Ensembles of independent 2D HLLC problems on one GPU (imogen_ensemble_*).

The workload of the batched-ensemble benchmark: `count` four-quadrant Riemann
problems on the unit square, N_k x N_k cells with N_k = round(nmin (nmax /
nmin)^u), u ~ U[0, 1] (log-uniform sizes), quadrant states rho, P ~ U[0.1, 2]
and v_x, v_y ~ U[-1, 1], gamma = 1.4, constant-extrapolation boundaries,
per-problem CFL 0.4. Odd-numbered problems may carry dust at dust-to-gas
ratio 0.5 (quadrant_dust, Ensemble.set_dust) coupled by the drag source.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .problems import conserved


def quadrant_dust(gas_state, seed, eps=0.5, dust_gamma=1.01, dust_cs=0.3):
    """Dust for a quadrant problem: rho_d = eps rho_g, its own quadrant
    velocities (U[-1, 1]), a cold gamma = 1.01 fluid with sound speed dust_cs."""
    rng = np.random.default_rng(seed)
    n = gas_state.shape[-1]
    ny = gas_state.shape[-2]
    rho = eps * gas_state[0]
    vx, vy = rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
    h, hy = n // 2, ny // 2
    q_idx = np.zeros((ny, n), dtype=int)
    q_idx[:hy, h:] = 1
    q_idx[hy:, :h] = 2
    q_idx[hy:, h:] = 3
    return conserved(rho, vx[q_idx], vy[q_idx], np.zeros((ny, n)), rho * dust_cs ** 2 / dust_gamma, dust_gamma)


def quadrant_problems(count=512, seed=0, nmin=64, nmax=512, gamma=1.4):
    rng = np.random.default_rng(seed)
    sizes = np.rint(nmin * (nmax / nmin) ** rng.uniform(0, 1, count)).astype(int)
    states = []
    for n in sizes:
        rho, P = rng.uniform(0.1, 2.0, 4), rng.uniform(0.1, 2.0, 4)
        vx, vy = rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4)
        h = n // 2
        q_idx = np.zeros((n, n), dtype=int)  # quadrant index: [y, x]
        q_idx[:h, h:] = 1
        q_idx[h:, :h] = 2
        q_idx[h:, h:] = 3
        states.append(conserved(rho[q_idx], vx[q_idx], vy[q_idx], np.zeros((n, n)), P[q_idx], gamma))
    return sizes, states


class Ensemble:
    def __init__(self, nx, ny, hx, hy, gamma=1.4, limiter=abi.LIMITER_OSPRE, cfl=0.4):
        self.nx = np.ascontiguousarray(nx, dtype=np.int32)
        self.ny = np.ascontiguousarray(ny, dtype=np.int32)
        hx = np.ascontiguousarray(hx, dtype=np.float64)
        hy = np.ascontiguousarray(hy, dtype=np.float64)
        self.count = len(self.nx)
        self.cells = int(np.sum(self.nx.astype(np.int64) * self.ny))
        h = ctypes.c_void_p(0)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        abi.call("imogen_ensemble_create", self.count, self.nx.ctypes.data_as(ip), self.ny.ctypes.data_as(ip),
                 hx.ctypes.data_as(dp), hy.ctypes.data_as(dp), float(gamma), int(limiter), float(cfl),
                 ctypes.byref(h))
        self.h = h
        self.iteration = 0

    def set_state(self, p, state):
        s = np.ascontiguousarray(state, dtype=np.float64)
        if s.shape != (5, self.ny[p], self.nx[p]):
            raise ValueError(f"problem {p} expects shape (5, {self.ny[p]}, {self.nx[p]})")
        abi.call("imogen_ensemble_set_state", self.h, int(p), s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def state(self, p):
        out = np.empty((5, self.ny[p], self.nx[p]))
        abi.call("imogen_ensemble_get_state", self.h, int(p), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def set_dust(self, has_dust, gas_thermo, dust_thermo, dust_gamma=1.01, method=abi.DRAG_LEARNED):
        """Problems with has_dust[p] carry a dust fluid coupled by the drag source."""
        flags = np.ascontiguousarray(np.asarray(has_dust, dtype=bool).astype(np.int32))
        if flags.shape != (self.count,):
            raise ValueError("one dust flag per problem")
        abi.call("imogen_ensemble_set_dust", self.h, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                 float(dust_gamma), ctypes.byref(gas_thermo), ctypes.byref(dust_thermo), int(method))
        self.dusty = flags.astype(bool)
        self.dust_cells = int(np.sum((self.nx.astype(np.int64) * self.ny)[self.dusty]))

    def set_dust_state(self, p, state):
        s = np.ascontiguousarray(state, dtype=np.float64)
        if s.shape != (5, self.ny[p], self.nx[p]):
            raise ValueError(f"problem {p} expects shape (5, {self.ny[p]}, {self.nx[p]})")
        abi.call("imogen_ensemble_set_dust_state", self.h, int(p), s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def dust_state(self, p):
        out = np.empty((5, self.ny[p], self.nx[p]))
        abi.call("imogen_ensemble_get_dust_state", self.h, int(p), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def run(self, iterations, graph=True):
        abi.call("imogen_ensemble_run", self.h, self.iteration, int(iterations), 1 if graph else 0)
        self.iteration += iterations

    def times(self):
        t = np.empty(self.count)
        abi.call("imogen_ensemble_times", self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return t

    def autotune(self, trials=2):
        chosen = (ctypes.c_int * 3)()
        abi.call("imogen_ensemble_autotune", self.h, int(trials), chosen)
        return tuple(chosen)

    def x_lanes(self):
        """(cells per lane, split) of the X sweep per line-length group
        (nx <= 64, 128, 192, 256, 368, 496, 512); (0, 0) for empty groups."""
        k = (ctypes.c_int * 7)()
        sp = (ctypes.c_int * 7)()
        abi.call("imogen_ensemble_x_lanes", self.h, k, sp)
        return tuple(zip(tuple(k), tuple(sp)))

    def free(self):
        if self.h:
            abi.call("imogen_ensemble_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
