"""The Imogen main loop on one GPU, for one or more fluids.

Restates the iteration of imogen.m:115-139 and the time-step rule of
time/TimeManager.m:124-186 on top of this library's entry points:

    dTime = CFL * min over fluids of cflTimestep(fluid)  (clamped so that
            t + 2 dTime does not pass tMax)
    source(1/2); fluidstep(+1); source(1); fluidstep(-1); source(1/2)
    t += 2 dTime

where source(f) is sources/src0001.m (two gas-dust drag calls over
f dTime / 2 each, then every fluid's boundary conditions) when a drag method
is set, and nothing otherwise.
"""
from __future__ import annotations

import ctypes

from . import abi
from . import gpuclass as gc


class Simulation:
    def __init__(self, fluids, methods, geometry, cfl=0.4, drag_method=None, limiter=abi.LIMITER_OSPRE,
                 t_max=None):
        if len(methods) != len(fluids):
            raise ValueError("one fluid method per fluid")
        if drag_method is not None and len(fluids) != 2:
            raise ValueError("the drag source couples exactly two fluids (gas, dust)")
        self.fluids = list(fluids)
        self.methods = [int(m) for m in methods]
        self.geometry = geometry
        self.cfl = float(cfl)
        self.drag_method = drag_method
        self.limiter = int(limiter)
        self.t_max = t_max
        self.time = 0.0
        self.iteration = 0

    def time_step(self):
        """TimeManager.update: the CFL-limited dTime over all fluids."""
        tau = min(gc.cflTimestepFused(f, self.geometry, m) for f, m in zip(self.fluids, self.methods))
        dt = self.cfl * tau
        if self.t_max is not None and self.time + 2 * dt > self.t_max:
            dt = 0.5 * (self.t_max - self.time)
        return dt

    def source(self, dt, fraction):
        if self.drag_method is not None:
            gc.src0001(self.fluids, self.geometry, dt, fraction, self.drag_method)

    def fluidstep(self, dt, direction):
        for f, m in zip(self.fluids, self.methods):
            gc.fluidstep([f], None, None, None, [dt, 1, direction, self.iteration, m], self.geometry,
                         limiter=self.limiter)

    def step(self):
        dt = self.time_step()
        self.source(dt, 0.5)
        self.fluidstep(dt, 1)
        self.source(dt, 1.0)
        self.fluidstep(dt, -1)
        self.source(dt, 0.5)
        self.time += 2 * dt
        self.iteration += 1
        return dt


class DeviceLoop:
    """Iterations of Simulation (with the drag source when drag_method is
    set) with dTime kept on the device and each iteration replayed from a HIP
    graph (imogen_timeloop_*). Bitwise the same results as Simulation.step()."""

    def __init__(self, fluids, methods, geometry, cfl=0.4, limiter=abi.LIMITER_OSPRE, t_max=None,
                 drag_method=None):
        if len(methods) != len(fluids):
            raise ValueError("one fluid method per fluid")
        if drag_method is not None and len(fluids) != 2:
            raise ValueError("the drag source couples exactly two fluids (gas, dust)")
        self.fluids = list(fluids)
        nf = len(fluids)
        arr = (abi.MGArray * (5 * nf))()
        params = (abi.FluidStepParams * nf)()
        gam = (ctypes.c_double * nf)()
        for i, (f, m) in enumerate(zip(fluids, methods)):
            a = f.arrays()
            for k in range(5):
                arr[5 * i + k] = a[k]
            params[i] = gc.step_params(0.0, 1, 0, int(m), geometry, f.gamma, f.MINMASS, f.Cisothermal, limiter)
            gam[i] = f.gamma
        g = geometry.params()
        h = ctypes.c_void_p(0)
        abi.call("imogen_timeloop_create", arr, nf, ctypes.byref(g), params, gam, float(cfl),
                 float(t_max) if t_max is not None else 0.0, ctypes.byref(h))
        self.h = h
        self.iteration = 0
        if drag_method is not None:
            abi.call("imogen_timeloop_set_drag", self.h, ctypes.byref(fluids[0].thermoDetails),
                     ctypes.byref(fluids[1].thermoDetails), int(drag_method))

    def run(self, iterations, graph=True):
        abi.call("imogen_timeloop_run", self.h, self.iteration, int(iterations), 1 if graph else 0)
        self.iteration += int(iterations)

    def clock(self):
        t, dt = ctypes.c_double(0.0), ctypes.c_double(0.0)
        abi.call("imogen_timeloop_clock", self.h, ctypes.byref(t), ctypes.byref(dt))
        return t.value, dt.value

    @property
    def time(self):
        return self.clock()[0]

    def free(self):
        if self.h:
            abi.call("imogen_timeloop_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
