"""The Imogen iteration run on the CPU oracle (test helper, no tests).

imogen.m:115-139 with TimeManager.m:124-186's time step: per iteration
    dTime = CFL * min over fluids of the CFL limit (clamped to t_max)
    source(1/2); fluidstep(+1) of every fluid; source(1); fluidstep(-1); source(1/2)
where source(f) is sources/src0001.m: two drag calls over f dTime / 2 each,
then every fluid's boundary conditions in the directions the grid has
(FluidManager.setBoundaries) -- only when a drag method is given.
"""
import numpy as np

from oracle import pyoracle as po


class OracleFluid:
    def __init__(self, q, gamma, method, modes, limiter=po.LIM_OSPRE, thermo=None):
        self.q = np.array(q, dtype=np.float64, copy=True)
        self.gamma = gamma
        self.method = method
        self.modes = tuple(modes)
        self.limiter = limiter
        self.thermo = thermo


def _dims(q):
    return q.shape[1:][::-1]  # (nx, ny, nz)


def oracle_iterations(fluids, h, iterations, cfl=0.4, first_iteration=0, t0=0.0, t_max=None, drag=None,
                      drag_tol=1e-12):
    """Advances the OracleFluids in place; drag = method code (fluids[0] gas,
    fluids[1] dust, their .thermo set). Returns the simulation time reached."""
    t = t0
    n = _dims(fluids[0].q)
    for k in range(iterations):
        it = first_iteration + k
        tau = min(po.cfl_timestep(f.q, po.soundspeed(f.q, f.gamma), h, f.method, n) for f in fluids)
        dt = cfl * tau
        if t_max is not None and t + 2 * dt > t_max:
            dt = 0.5 * (t_max - t)

        def source(frac):
            if drag is None:
                return
            g, d = fluids[0], fluids[1]
            for _ in range(2):
                po.drag(g.q, d.q, g.thermo, d.thermo, dt * frac / 2, drag, tol=drag_tol)
            for f in fluids:
                for direction in (1, 2, 3):
                    if direction == 1 or n[direction - 1] > 1:
                        po.set_fluid_boundary(f.q, direction, f.modes)

        source(0.5)
        for sgn in (1, -1):
            for f in fluids:
                po.fluidstep(f.q, h, dt, sgn, it, f.gamma, f.method, f.limiter, modes=f.modes)
            source(1.0 if sgn > 0 else 0.5)
        t += 2 * dt
    return t


def rel_l1(a, b):
    """Relative L1 difference per slab (5 values)."""
    return np.array([np.abs(a[s] - b[s]).sum() / max(np.abs(b[s]).sum(), 1e-300) for s in range(5)])
