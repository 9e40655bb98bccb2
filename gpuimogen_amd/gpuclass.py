"""Host-side mirror of the reference's MATLAB layer over libimogen_hip.so.

The reference drives its GPU code from MATLAB through MEX gateways taking
GPU_Type / FluidManager / GeometryManager objects. MATLAB is not part of this
build, so this module restates those objects and call forms in Python, on top
of the C ABI (include/imogen_abi.h), with the same argument meaning:

    fluidstep(fluids, bx, by, bz, [dt, purehydro, +-1, stepNumber, method], geometry)
        gpuclass/flux_ML_iface.cu:37-162        (called at imogen.m:122,125)
    tau = cflTimestep(fluid, soundspeed, geometry, method)
        gpuclass/cflTimestep.cu:64-103          (called at time/TimeManager.m:150)
    cs = cudaSoundspeed(mass, ener, momX, momY, momZ, gamma)
        gpuclass/cudaSoundspeed.cu:54-100       (called at time/TimeManager.m:143)
    cudaStatics(array, 8, geometry, direction)
        gpuclass/cudaStatics.cu:85-104          (ImogenArray.m:266-268)

Host arrays are numpy float64 with x fastest: a scalar field is shaped
(nz, ny, nx) in C order, a fluid state (5, nz, ny, nx) with slabs
(rho, E, px, py, pz), the same memory order as the device slabs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .abi import call

HALO = 4  # FLUID_METHOD_HALO_SIZE under USE_SSPRK (gpuclass/fluidMethod.h:13)


def _dims_of(a):
    a = np.asarray(a)
    if a.ndim == 1:
        return (a.shape[0], 1, 1)
    if a.ndim == 2:
        return (a.shape[1], a.shape[0], 1)
    if a.ndim == 3:
        return (a.shape[2], a.shape[1], a.shape[0])
    raise ValueError("arrays must be 1-, 2- or 3-dimensional")


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class GPU_Type:
    """A device array addressed by an int64 tag (gpuclass/GPU_Type.m)."""

    def __init__(self, host=None, dims=None, halo=HALO, tag=None, owner=True):
        self.owner = owner
        if tag is not None:
            self.tag = np.array(tag, dtype=np.int64)
            return
        if host is not None:
            host = np.ascontiguousarray(host, dtype=np.float64)
            dims = _dims_of(host)
        if dims is None:
            raise ValueError("need host data or dims")
        self.tag = np.zeros(abi.TAG_LEN_1GPU, dtype=np.int64)
        d = (ctypes.c_int * 3)(*dims)
        call("GPU_upload", _dptr(host) if host is not None else None, d, halo, abi.PARTITION_X,
             self._tp())

    def _tp(self, t=None):
        t = self.tag if t is None else t
        return t.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))

    @property
    def dims(self):
        return tuple(int(x) for x in self.tag[:3])

    @property
    def numSlabs(self):
        return int(self.tag[3])

    def mgarray(self):
        m = abi.MGArray()
        call("deserializeTagToMGArray", self._tp(), ctypes.byref(m))
        return m

    def createSlabs(self, n):
        """GPU_Type.createSlabs (GPU_Type.m:166-169): realloc to n slabs, slab 0 kept."""
        nt = np.zeros_like(self.tag)
        call("GPU_makeslab", self._tp(), n, self._tp(nt))
        self.tag = nt

    def slab(self, k):
        vt = np.zeros_like(self.tag)
        call("GPU_getslab", self._tp(), k, self._tp(vt))
        return GPU_Type(tag=vt, owner=False)

    def setslab(self, k, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        vt = np.zeros_like(self.tag)
        call("GPU_setslab", self._tp(), k, _dptr(host), self._tp(vt))
        return GPU_Type(tag=vt, owner=False)

    @property
    def array(self):
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx), dtype=np.float64)
        call("GPU_download", self._tp(), _dptr(out))
        return out

    def free(self):
        if self.owner and self.tag is not None and self.tag[3] > 0:
            call("GPU_free", self._tp())
        self.tag = None

    def __del__(self):
        try:
            if getattr(self, "owner", False) and getattr(self, "tag", None) is not None:
                self.free()
        except Exception:
            pass


class GeometryManager:
    """Cartesian grid description (gpuclass/GeometryManager.m, accessed by
    accessMatlabGeometryClass, gpuclass/cudaCommonML.cu:453)."""

    def __init__(self, global_rez, h):
        self.globalDomainRez = tuple(int(x) for x in global_rez) + (1,) * (3 - len(global_rez))
        self.h = tuple(float(x) for x in h) + (1.0,) * (3 - len(h))

    def params(self):
        g = abi.GeometryParams()
        g.shape = abi.SQUARE
        for i in range(3):
            g.h[i] = self.h[i]
            g.globalRez[i] = self.globalDomainRez[i]
            g.localRez[i] = self.globalDomainRez[i]
        return g


class FluidManager:
    """One fluid: a 5-slab DataHolder (rho, E, px, py, pz) plus thermodynamics
    and boundary modes (fluid/FluidManager.m, save/uploadDataArrays.m:93-111)."""

    def __init__(self, state, gamma, minmass=1e-8, bc_modes=(abi.circular,) * 6, cisothermal=-1.0,
                 thermoDetails=None):
        state = np.ascontiguousarray(state, dtype=np.float64)
        if state.ndim != 4 or state.shape[0] != 5:
            raise ValueError("fluid state must be shaped (5, nz, ny, nx)")
        self.gamma = float(gamma)
        self.MINMASS = float(minmass)
        self.Cisothermal = float(cisothermal)
        # microphysics for the drag source (FluidManager.thermoDetails)
        self.thermoDetails = thermoDetails if thermoDetails is not None else abi.ThermoDetails(
            self.gamma, 1.0, self.Cisothermal, 1.0, 0.0, 0.0, 1.0, 0.0)
        self.bcModes = tuple(int(m) for m in bc_modes)
        self.DataHolder = GPU_Type(state[0])
        self.DataHolder.createSlabs(5)
        self.mass = self.DataHolder.slab(0)
        self.ener = self.DataHolder.setslab(1, state[1])
        self.mom = [self.DataHolder.setslab(2 + i, state[2 + i]) for i in range(3)]

    def arrays(self):
        """The 5 MGArrays of the canister (MGA_accessFluidCanister, cudaCommonML.cu:400)."""
        arr = (abi.MGArray * 5)()
        for i, g in enumerate([self.mass, self.ener] + self.mom):
            m = g.mgarray()
            m.vectorComponent = 0 if i < 2 else i - 1
            for k in range(6):
                m.boundaryConditions.mode[k] = self.bcModes[k]
            arr[i] = m
        return arr

    @property
    def state(self):
        return np.stack([self.mass.array, self.ener.array] + [m.array for m in self.mom])

    def set_state(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        self.DataHolder.setslab(0, state[0])
        for i in range(1, 5):
            self.DataHolder.setslab(i, state[i])

    def free(self):
        self.DataHolder.free()


def step_params(dt, step_direction, step_number, method, geometry, gamma, minmass, cisothermal=-1.0,
                limiter=abi.LIMITER_OSPRE):
    p = abi.FluidStepParams()
    p.dt = dt
    p.onlyHydro = 1
    p.thermoGamma = gamma if cisothermal == -1 else 2.0  # flux_ML_iface.cu:128-133
    p.minimumRho = minmass
    p.Cisothermal = cisothermal
    p.stepMethod = method
    p.stepDirection = step_direction
    p.stepNumber = step_number
    p.geometry = geometry.params()
    p.slopeLimiter = limiter
    return p


def fluidstep(fluids, bx, by, bz, params, geometry, limiter=abi.LIMITER_OSPRE):
    """fluidstep(FluidManager[], bx, by, bz, [dt, purehydro, +-1, stepnum, method], geometry)
    (gpuclass/flux_ML_iface.cu:37-162)."""
    if len(params) != 5:
        raise ValueError("Must rx 5 parameters in params vector: [dt, purehydro?, order, step #, step method]")
    dt, purehydro, sweep_dir, stepnum, method = params
    if int(purehydro) != 1:
        raise NotImplementedError("MHD is not built")
    topo = abi.ParallelTopology()
    for i in range(3):
        topo.nproc[i] = 1
    for f in fluids:
        arr = f.arrays()
        p = step_params(float(dt), int(sweep_dir), int(stepnum), int(method), geometry, f.gamma,
                        f.MINMASS, f.Cisothermal, limiter)
        call("performFluidUpdate_3D", arr, ctypes.byref(topo), p, None)


def cudaFluidStep(params, fluid, geometry, limiter=abi.LIMITER_OSPRE):
    """cudaFluidStep([dt, purehydro, gamma, MINMASS, method, dir (, Ciso)], rho, E, px, py, pz,
    bx, by, bz, geometry): one directional update (gpuclass/cudaFluidStep.cu:160-227),
    as called by fluid/relaxingFluid.m:27."""
    if len(params) not in (6, 7):
        raise ValueError("need [dt, purehydro, gamma, MINMASS, method, direction (, C_iso)]")
    dt, purehydro, gamma, rhomin, method, direction = params[:6]
    ciso = params[6] if len(params) == 7 else -1.0
    if len(params) == 7:
        gamma = 2.0
    p = step_params(float(dt), int(direction), 0, int(method), geometry, float(gamma), float(rhomin),
                    float(ciso), limiter)
    p.thermoGamma = float(gamma)
    topo = abi.ParallelTopology()
    for i in range(3):
        topo.nproc[i] = 1
    call("performFluidUpdate_1D", fluid.arrays(), p, ctypes.byref(topo), None)


def cudaArrayRotateB(array, code):
    """cudaArrayRotateB(tag, code) in place (gpuclass/cudaArrayRotateB.cu:56-88):
    flipArrayIndices on one array, then its tag is rewritten (dims and
    permtag). code 2 = X<->Y, 3 = X<->Z, 4 = Y<->Z, 5 = XYZ -> YZX, 6 = XYZ -> ZXY."""
    m = array.mgarray()
    call("flipArrayIndices", ctypes.byref(m), None, 1, int(code), None, None)
    abi.lib().serializeMGArrayToTag(ctypes.byref(m), array._tp())


def cudaHaloExchange(array, direction, topology=None):
    """cudaHaloExchange(tag, dir, topology, circ) (cudaHaloExchange.cu:113-133):
    exchange_MPI_Halos on one array; nothing to exchange on one rank."""
    m = array.mgarray()
    topo = topology if topology is not None else abi.ParallelTopology()
    if topology is None:
        for i in range(3):
            topo.nproc[i] = 1
    call("exchange_MPI_Halos", ctypes.byref(m), 1, ctypes.byref(topo), int(direction))


def cudaStatics(array, geometry, direction, vector_component=0, bc_modes=None):
    """cudaStatics(array, 8, geometry, dir) (cudaStatics.cu:85-104): both faces
    of one array in physical direction dir (mapped to the memory axis that
    holds it)."""
    m = array.mgarray()
    m.vectorComponent = vector_component
    if bc_modes is not None:
        for k in range(6):
            m.boundaryConditions.mode[k] = int(bc_modes[k])
    g = geometry.params()
    for side in (0, 1):
        call("setArrayBoundaryConditions", ctypes.byref(m), ctypes.byref(g), int(direction), side)


def xchgIndices(fluid, code):
    """flux/flux.m:xchgIndices: rotate the 5 arrays of a fluid (code 0: nothing)."""
    if code == 0:
        return
    for g in [fluid.mass, fluid.ener] + fluid.mom:
        cudaArrayRotateB(g, code)


def relaxingFluid(fluid, geometry, direction, dt, method, limiter=abi.LIMITER_OSPRE):
    """fluid/relaxingFluid.m:25-31: cudaFluidStep in physical direction X on
    the (rotated) arrays, then cudaStatics of each array in X."""
    p = [dt, 1, fluid.gamma, fluid.MINMASS, method, direction]
    if fluid.Cisothermal != -1:
        p.append(fluid.Cisothermal)
    cudaFluidStep(p, fluid, geometry, limiter=limiter)
    for i, g in enumerate([fluid.mass, fluid.ener] + fluid.mom):
        cudaStatics(g, geometry, direction, vector_component=0 if i < 2 else i - 1, bc_modes=fluid.bcModes)


# flux/flux.m:29-40
_PREPERM = [0, 2, 0, 2, 3, 3]
_FLUXCALL = [[1, 2, 1, 2, 3, 3], [3, 1, 2, 3, 1, 2], [2, 3, 3, 1, 2, 1]]
_PERMCALL = [[3, 2, 2, 3, 3, 2], [2, 3, 3, 5, 2, 6], [6, 3, 5, 0, 2, 0]]


def flux(fluids, geometry, order, iteration, dt, method, limiter=abi.LIMITER_OSPRE):
    """flux/flux.m: the MATLAB-side fluid step of one direction order (+1
    forward, -1 backward) that rotates the arrays so every sweep is along
    memory x: xchgIndices(preperm), then per position n relaxingFluid of
    fluxcall(n) (directions of more than 2 cells), the (one-rank) halo
    exchange, xchgIndices(permcall(n)). The tables end every sequence in XYZ
    order. 2D runs take sweep = mod(iteration + (order < 0), 2) + 1 as flux.m
    does; 1D runs one X sweep."""
    dims = geometry.globalDomainRez
    if sum(1 for d in dims if d > 2) == 1:
        for f in fluids:
            relaxingFluid(f, geometry, 1, dt, method, limiter)
        return
    if dims[2] > 2:
        sweep = (iteration + 3 * (order > 0)) % 6
    else:
        sweep = (iteration + (order < 0)) % 2
    for f in fluids:
        xchgIndices(f, _PREPERM[sweep])
        for n in range(3):
            d = _FLUXCALL[n][sweep]
            if dims[d - 1] > 2:
                relaxingFluid(f, geometry, d, dt, method, limiter)
                for g in [f.mass, f.ener] + f.mom:
                    cudaHaloExchange(g, d)
            xchgIndices(f, _PERMCALL[n][sweep])


def cudaSoundspeed(fluid, gamma):
    """cs = cudaSoundspeed(mass, ener, momX, momY, momZ, gamma) (cudaSoundspeed.cu:54-100)."""
    arr = fluid.arrays()
    cs = GPU_Type(dims=fluid.mass.dims)
    csm = cs.mgarray()
    call("calculateSoundspeed", arr, None, ctypes.byref(csm), float(gamma))
    return cs


def cflTimestep(fluid, soundspeed, geometry, method):
    """tau = cflTimestep(FluidManager, soundspeed, GeometryManager, method) (cflTimestep.cu:64-103)."""
    arr = fluid.arrays()
    csm = soundspeed.mgarray()
    g = geometry.params()
    rez = (ctypes.c_double * 3)(*geometry.globalDomainRez)
    tau = ctypes.c_double(0.0)
    call("computeLocalCFLTimestep", arr, ctypes.byref(csm), ctypes.byref(g), int(method), rez,
         ctypes.byref(tau))
    return tau.value


def cflTimestepFused(fluid, geometry, method):
    arr = fluid.arrays()
    g = geometry.params()
    rez = (ctypes.c_double * 3)(*geometry.globalDomainRez)
    tau = ctypes.c_double(0.0)
    call("imogen_cflTimestepFused", arr, ctypes.byref(g), int(method), fluid.gamma, rez, ctypes.byref(tau))
    return tau.value


def setBoundaries(fluid, geometry, direction):
    """FluidManager.setBoundaries(direction) -> setFluidBoundary (cudaStatics.cu:112-196)."""
    arr = fluid.arrays()
    g = geometry.params()
    call("setFluidBoundary", arr, ctypes.byref(g), int(direction))


def synchronize():
    call("imogen_synchronize")


def cudaSource2FluidDrag(fluids, geometry, params):
    """cudaSource2FluidDrag(FluidManager[2], geometry, [dt, method])
    (gpuclass/cudaSource2FluidDrag.cu:125-180): fluids[0] is the gas,
    fluids[1] the dust."""
    if len(fluids) != 2 or len(params) != 2:
        raise ValueError("need cudaSource2FluidDrag(FluidManager[2], geometry, [dt, solverMethod])")
    g = geometry.params()
    call("sourcefunction_2FluidDrag", fluids[0].arrays(), fluids[1].arrays(), ctypes.byref(g),
         ctypes.byref(fluids[0].thermoDetails), ctypes.byref(fluids[1].thermoDetails), float(params[0]),
         int(params[1]))


def sourceStep(fluids, geometry, dTime, params, multifluidDragMethod=abi.DRAG_LOGTRAP2, radiation=None):
    """sourceStep(run, fluid, bx, by, bz, xyvector, [order, step, method, tFraction])
    (gpuclass/sourceStep.cu:71-193, the MEX form of performSourceFunctions):
    dt = dTime * tFraction, srcType = cylindrical + 2 rotatingFrame + 8 (more
    than one fluid) -- no gravity potential here -- then
    performSourceFunctions (sourceStep.cu:200-226). With two fluids that is the
    drag source over the whole dt with multifluidDragMethod (fluids[0] gas,
    fluids[1] dust). radiation: (exponent, strength) or None; any strength
    raises ERROR_NOIMPLEMENT, as do the cylindrical and rotating-frame types."""
    if len(params) != 4:
        raise ValueError("need [order, step #, step method, tFraction]")
    order, step, method, tFraction = params
    dt = float(dTime) * float(tFraction)
    n = len(fluids)
    gf = (abi.GridFluid * n)()
    for i, f in enumerate(fluids):
        arr = f.arrays()
        gf[i].DataHolder = arr[0]
        for k in range(5):
            gf[i].data[k] = arr[k]
            gf[i].bcond[k] = arr[k].boundaryConditions
        gf[i].thermo = f.thermoDetails
        gf[i].rhoMin = f.MINMASS
    f0 = fluids[0]
    fsp = step_params(dt, int(order), int(step), int(method), geometry, f0.gamma, f0.MINMASS, f0.Cisothermal)
    fsp.multifluidDragMethod = int(multifluidDragMethod) if n > 1 else 0
    g = geometry.params()
    cyl = g.shape in (abi.CYLINDRICAL, abi.RZCYLINDRICAL)
    srcType = 1 * int(cyl) + 2 * int(g.frameOmega != 0) + 8 * int(n > 1)
    rad = abi.ParametricRadiation()
    if radiation is not None:
        rad.exponent = float(radiation[0])
        rad.minTemperature = 1.05  # sourceStep.cu:122
        rad.prefactor = float(radiation[1]) * dt
    topo = abi.ParallelTopology()
    for i in range(3):
        topo.nproc[i] = 1
    call("performSourceFunctions", srcType, gf, n, fsp, ctypes.byref(topo), None, ctypes.byref(rad), None)


def src0001(fluids, geometry, dTime, tFraction, dragMethod):
    """sources/src0001.m: drag over dTime*tFraction as two half steps, then the
    boundary conditions of every fluid in every direction."""
    dt = dTime * tFraction
    cudaSource2FluidDrag(fluids, geometry, [dt / 2, dragMethod])
    cudaSource2FluidDrag(fluids, geometry, [dt / 2, dragMethod])
    for f in fluids:
        for d in (1, 2, 3):
            if geometry.globalDomainRez[d - 1] > 1 or d == 1:
                setBoundaries(f, geometry, d)
