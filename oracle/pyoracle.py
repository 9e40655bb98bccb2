"""ctypes binding of oracle/liboracle.so (test infrastructure only).

The functions take numpy float64 arrays shaped (5, nz, ny, nx) -- the
5-slab fluid canister (rho, E, px, py, pz) with x fastest -- and update them
in place, exactly as the CPU restatement in imogen_oracle.c does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

LIM_OSPRE, LIM_VANLEER, LIM_MINMOD, LIM_SUPERBEE, LIM_VANALBADA, LIM_MC, LIM_ZERO = range(7)
HLL, HLLC, XINJIN = 1, 2, 3
BC_CIRCULAR, BC_MIRROR, BC_WALL, BC_STATIONARY, BC_CONST, BC_LINEAR, BC_OUTFLOW, BC_FREEBALANCE = range(8)


class _Params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("rhomin", ctypes.c_double),
                ("ciso", ctypes.c_double), ("limiter", ctypes.c_int)]


class _Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("pitch", ctypes.c_long)]


class Thermo(ctypes.Structure):
    """ThermoDetails (gpuclass/cudaCommon.h:146-158)."""
    _fields_ = [("gamma", ctypes.c_double), ("m", ctypes.c_double), ("Cisothermal", ctypes.c_double),
                ("kBolt", ctypes.c_double), ("mu0", ctypes.c_double), ("muTindex", ctypes.c_double),
                ("sigma0", ctypes.c_double), ("sigmaTindex", ctypes.c_double)]


DRAG_EMP, DRAG_RK4, DRAG_ETDRK1, DRAG_LOGTRAP2, DRAG_LOGTRAP3, DRAG_EXACT = 0, 1, 2, 4, 5, 6

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = ctypes.CDLL(LIBPATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_fluid_update_1d.argtypes = [dp, ctypes.POINTER(_Grid), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.POINTER(_Params)]
        L.orc_fluid_update_1d.restype = ctypes.c_int
        L.orc_xj_predict_1d.argtypes = [dp, ctypes.POINTER(_Grid), ctypes.c_int, ctypes.c_double,
                                        ctypes.POINTER(_Params)]
        L.orc_xj_predict_1d.restype = ctypes.c_int
        L.orc_set_fluid_boundary.argtypes = [dp, ctypes.POINTER(_Grid), ctypes.c_int, ip, ctypes.c_int]
        L.orc_freebalance_face.argtypes = [dp, ctypes.POINTER(_Grid), dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int]
        L.orc_static_cells.argtypes = [dp, dp, ctypes.c_int, ctypes.c_long]
        L.orc_static_cells.restype = None
        L.orc_set_fluid_boundary.restype = ctypes.c_int
        L.orc_fluid_boundary_swept.argtypes = [dp, ctypes.POINTER(_Grid), ctypes.c_int, ip, ctypes.c_int]
        L.orc_fluid_boundary_swept.restype = ctypes.c_int
        L.orc_fluidstep.argtypes = [dp, ctypes.POINTER(_Grid), dp, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Params), ip, ctypes.c_int]
        L.orc_fluidstep.restype = ctypes.c_int
        L.orc_sweep_order.argtypes = [ctypes.c_int, ip]
        L.orc_soundspeed.argtypes = [dp, ctypes.POINTER(_Grid), ctypes.c_double, dp]
        L.orc_cfl_timestep.argtypes = [dp, dp, ctypes.POINTER(_Grid), dp, ctypes.c_int, ip]
        L.orc_cfl_timestep.restype = ctypes.c_double
        L.orc_slope_limiter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_slope_limiter.restype = ctypes.c_double
        L.orc_drag_params.argtypes = [ctypes.POINTER(Thermo), ctypes.POINTER(Thermo), dp]
        L.orc_drag_k.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_drag_k.restype = ctypes.c_double
        L.orc_drag_rk_magdv.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int]
        L.orc_drag_rk_magdv.restype = ctypes.c_double
        L.orc_drag_exact_ratio.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ip]
        L.orc_drag_exact_ratio.restype = ctypes.c_double
        L.orc_drag.argtypes = [dp, dp, ctypes.POINTER(_Grid), ctypes.POINTER(Thermo), ctypes.POINTER(Thermo),
                               ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_long)]
        L.orc_drag.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _grid(q):
    if q.dtype != np.float64 or not q.flags.c_contiguous or q.ndim != 4 or q.shape[0] != 5:
        raise ValueError("fluid must be C-contiguous float64 shaped (5, nz, ny, nx)")
    _, nz, ny, nx = q.shape
    return _Grid(nx, ny, nz, nx * ny * nz)


def _params(gamma, rhomin=1e-8, ciso=-1.0, limiter=LIM_OSPRE):
    return _Params(gamma, rhomin, ciso, limiter)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def fluid_update_1d(q, direction, lam, gamma, method=HLLC, limiter=LIM_OSPRE, rhomin=1e-8, ciso=-1.0):
    g = _grid(q)
    p = _params(gamma, rhomin, ciso, limiter)
    rc = lib().orc_fluid_update_1d(_dptr(q), ctypes.byref(g), direction, method, lam, ctypes.byref(p))
    if rc:
        raise RuntimeError(f"orc_fluid_update_1d returned {rc}")


def xj_predict_1d(q, direction, lam, gamma, rhomin=1e-8):
    """Xin-Jin predictor only (freeze + U = floor(Q - lam/4 dF(Q)))."""
    g = _grid(q)
    p = _params(gamma, rhomin, -1.0, LIM_OSPRE)
    rc = lib().orc_xj_predict_1d(_dptr(q), ctypes.byref(g), direction, lam, ctypes.byref(p))
    if rc:
        raise RuntimeError(f"orc_xj_predict_1d returned {rc}")


def _modes(modes):
    m = np.asarray(modes, dtype=np.int32)
    if m.shape != (6,):
        raise ValueError("modes must be 6 ints [-x,+x,-y,+y,-z,+z]")
    return m


def set_fluid_boundary(q, direction, modes, depth=4):
    g = _grid(q)
    m = _modes(modes)
    rc = lib().orc_set_fluid_boundary(_dptr(q), ctypes.byref(g), direction,
                                      m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), depth)
    if rc:
        raise RuntimeError(f"orc_set_fluid_boundary returned {rc}")


def freebalance_face(q, phi, direction, side, gamma, rinner=0.0, h0=1.0, frame_omega=0.0, depth=4):
    """Freebalance ghosts of one face of q (5, nz, ny, nx) in place; phi (nz, ny, nx)."""
    g = _grid(q)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    assert ph.shape == q.shape[1:]
    rc = lib().orc_freebalance_face(_dptr(q), ctypes.byref(g), _dptr(ph), direction, side, depth, gamma, rinner,
                                    h0, frame_omega, q.shape[3])
    if rc:
        raise RuntimeError(f"orc_freebalance_face returned {rc}")


def static_cells(phi, table):
    """Static-cell fade of one array in place: table (3, n) = cell index, f0, c."""
    t = np.ascontiguousarray(table, dtype=np.float64)
    assert phi.flags.c_contiguous and phi.dtype == np.float64
    lib().orc_static_cells(_dptr(phi), _dptr(t), t.shape[1], t.shape[1])


def fluid_boundary_swept(q, direction, modes, depth=4):
    """The boundaries the fluid step sets after sweeping `direction`."""
    g = _grid(q)
    m = _modes(modes)
    rc = lib().orc_fluid_boundary_swept(_dptr(q), ctypes.byref(g), direction,
                                        m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), depth)
    if rc:
        raise RuntimeError(f"orc_fluid_boundary_swept returned {rc}")


def fluidstep(q, h, dt, step_direction, step_number, gamma, method=HLLC, limiter=LIM_OSPRE,
              modes=(BC_CIRCULAR,) * 6, depth=4, rhomin=1e-8, ciso=-1.0):
    g = _grid(q)
    p = _params(gamma, rhomin, ciso, limiter)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    m = _modes(modes)
    rc = lib().orc_fluidstep(_dptr(q), ctypes.byref(g), _dptr(hh), dt, step_direction, step_number,
                             method, ctypes.byref(p), m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), depth)
    if rc:
        raise RuntimeError(f"orc_fluidstep returned {rc}")


def sweep_order(sweep):
    out = (ctypes.c_int * 3)()
    lib().orc_sweep_order(sweep, out)
    return tuple(out)


def soundspeed(q, gamma):
    g = _grid(q)
    cs = np.empty(q.shape[1:], dtype=np.float64)
    lib().orc_soundspeed(_dptr(q), ctypes.byref(g), gamma, _dptr(cs))
    return cs


def cfl_timestep(q, cs, h, method, global_rez=None):
    g = _grid(q)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    if global_rez is None:
        global_rez = (q.shape[3], q.shape[2], q.shape[1])
    gr = np.asarray(global_rez, dtype=np.int32)
    cs = np.ascontiguousarray(cs, dtype=np.float64)
    return lib().orc_cfl_timestep(_dptr(q), _dptr(cs), ctypes.byref(g), _dptr(hh), method,
                                  gr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))


def slope_limiter(which, a, b):
    return lib().orc_slope_limiter(which, a, b)


def drag_params(tgas, tdust):
    P = np.zeros(9)
    lib().orc_drag_params(ctypes.byref(tgas), ctypes.byref(tdust), _dptr(P))
    return P


def drag_k(P, rhog, rhod, magdv, u):
    P = np.ascontiguousarray(P, dtype=np.float64)
    return lib().orc_drag_k(_dptr(P), rhog, rhod, magdv, u)


def drag_rk_magdv(P, rhog, rhod, s0, u0, dt, method):
    """|dv| after dt by the EMP (method 0) or RK4 (method 1) integrator."""
    return lib().orc_drag_rk_magdv(_dptr(P), rhog, rhod, s0, u0, dt, method)


def drag_exact_ratio(P, rhog, rhod, s0, u0, t, tol=1e-12):
    P = np.ascontiguousarray(P, dtype=np.float64)
    steps = ctypes.c_int(0)
    r = lib().orc_drag_exact_ratio(_dptr(P), rhog, rhod, s0, u0, t, tol, ctypes.byref(steps))
    return r, steps.value


DRAG_INVALID = -(1 << 30)  # drag_exact_ratio's step count for a cell it cannot integrate


class DragStateError(RuntimeError):
    """The exact integrator met cells it cannot integrate (every cell is still updated)."""


def drag(gas, dust, tgas, tdust, dt, method=DRAG_LOGTRAP2, tol=1e-12):
    """One drag source call (in place on both 5-slab states)."""
    g = _grid(gas)
    if dust.shape != gas.shape or not dust.flags.c_contiguous or dust.dtype != np.float64:
        raise ValueError("dust must match the gas array")
    work = ctypes.c_long(0)
    rc = lib().orc_drag(_dptr(gas), _dptr(dust), ctypes.byref(g), ctypes.byref(tgas), ctypes.byref(tdust), dt,
                        method, tol, ctypes.byref(work))
    if rc == -101:
        raise DragStateError("orc_drag: invalid cells or an exhausted step budget")
    if rc:
        raise RuntimeError(f"orc_drag returned {rc}")
    return work.value
