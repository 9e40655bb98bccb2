"""ctypes mirror of include/imogen_abi.h and the loader of libimogen_hip.so.

The structures below must stay field-for-field identical to the C header;
tests/test_abi.py compiles a probe against the header and compares every
size and offset.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# IMOGEN_HIP_LIB selects another build of the same library (e.g. a variant under test)
LIBPATH = os.environ.get("IMOGEN_HIP_LIB") or os.path.join(HERE, "libimogen_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "imogen_abi.h")

# status codes (gpuclass/imogenChecks.h:12-24)
SUCCESSFUL = 0
ERROR_INVALID_ARGS = -100
ERROR_CRASH = -101
ERROR_NULL_POINTER = -102
ERROR_GET_GPUTAG_FAILED = -103
ERROR_DESERIALIZE_GPUTAG_FAILED = -104
ERROR_NOIMPLEMENT = -105
ERROR_NOMEM = -106
ERROR_LIBFAILED = -107
ERROR_NOFILE = -108
ERROR_FILE_IO = -109
ERROR_CUDA_BLEW_UP = -1000

MAX_GPUS_USED = 4
GPU_TAG_LENGTH = 11
TAG_LEN_1GPU = GPU_TAG_LENGTH + 2
PARTITION_X, PARTITION_Y, PARTITION_Z = 1, 2, 3

# BCModeTypes
circular, mirror, wall, stationary, extrapConstant, extrapLinear, outflow, freebalance = range(8)
# geometryType_t
SQUARE, CYLINDRICAL, RZSQUARE, RZCYLINDRICAL = range(4)
# FluidMethods
METHOD_HLL, METHOD_HLLC, METHOD_XINJIN = 1, 2, 3
# SlopeLimiters
(LIMITER_OSPRE, LIMITER_VANLEER, LIMITER_MINMOD, LIMITER_SUPERBEE, LIMITER_VANALBADA, LIMITER_MC,
 LIMITER_ZERO) = range(7)
# DragMethods
(DRAG_EMP, DRAG_RK4, DRAG_ETDRK1, DRAG_ETDRK2, DRAG_LOGTRAP2, DRAG_LOGTRAP3, DRAG_EXACT,
 DRAG_LEARNED) = range(8)
IMOGEN_MLP_NPARAMS = 4739

c_int, c_double, c_int64, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_void_p


class BCSettings(ctypes.Structure):
    _fields_ = [("mode", c_int * 6), ("externalData", c_void_p), ("extIndex", c_int),
                ("staticCells", ctypes.POINTER(c_double) * MAX_GPUS_USED), ("nStatics", c_int * 4)]


class MGArray(ctypes.Structure):
    _fields_ = [("dim", c_int * 3), ("numel", c_int64), ("numSlabs", c_int),
                ("slabPitch", c_int64 * MAX_GPUS_USED), ("haloSize", c_int), ("partitionDir", c_int),
                ("addExteriorHalo", c_int), ("permtag", c_int), ("currentPermutation", c_int * 3),
                ("nGPUs", c_int), ("deviceID", c_int * MAX_GPUS_USED),
                ("devicePtr", c_void_p * MAX_GPUS_USED), ("partNumel", c_int * MAX_GPUS_USED),
                ("mpiCircularBoundaryBits", c_int), ("boundaryConditions", BCSettings),
                ("vectorComponent", c_int)]


class ImogenBoundaryData(ctypes.Structure):
    """What BCSettings.externalData points to: the freebalance potential and
    adiabatic index, and an array's static-cell table."""
    _fields_ = [("gravpot", ctypes.POINTER(MGArray)), ("gamma", c_double), ("statics", ctypes.POINTER(MGArray)),
                ("compOffset", c_double * (12 * MAX_GPUS_USED))]


class GeometryParams(ctypes.Structure):
    _fields_ = [("shape", c_int), ("h", c_double * 3), ("x0", c_double), ("y0", c_double),
                ("z0", c_double), ("Rinner", c_double), ("frameRotateCenter", c_double * 3),
                ("frameOmega", c_double), ("globalRez", c_int * 3), ("localRez", c_int * 3),
                ("gridAffine", c_int * 3), ("XYVector", c_void_p)]


class ThermoDetails(ctypes.Structure):
    _fields_ = [("gamma", c_double), ("m", c_double), ("Cisothermal", c_double), ("kBolt", c_double),
                ("mu0", c_double), ("muTindex", c_double), ("sigma0", c_double),
                ("sigmaTindex", c_double)]


class FluidStepParams(ctypes.Structure):
    _fields_ = [("dt", c_double), ("cflPrefactor", c_double), ("onlyHydro", c_int),
                ("thermoGamma", c_double), ("minimumRho", c_double), ("Cisothermal", c_double),
                ("stepMethod", c_int), ("stepDirection", c_int), ("stepNumber", c_int),
                ("geometry", GeometryParams), ("multifluidDragMethod", c_int), ("slopeLimiter", c_int)]


class ParallelTopology(ctypes.Structure):
    _fields_ = [("ndim", c_int), ("comm", c_int), ("nproc", c_int * 3), ("coord", c_int * 3),
                ("neighbor_left", c_int * 3), ("neighbor_right", c_int * 3), ("dimcomm", c_int * 3)]


class DragModelNorm(ctypes.Structure):
    _fields_ = [("fmean", ctypes.c_float * 5), ("fstd", ctypes.c_float * 5), ("tmean", ctypes.c_float * 3),
                ("tstd", ctypes.c_float * 3), ("flo", ctypes.c_float * 5), ("fhi", ctypes.c_float * 5),
                ("residTol", ctypes.c_float)]


class GravityData(ctypes.Structure):
    _fields_ = [("phi", ctypes.POINTER(MGArray)), ("spaceOrder", c_int), ("timeOrder", c_int)]


class GridFluid(ctypes.Structure):
    _fields_ = [("DataHolder", MGArray), ("data", MGArray * 5), ("thermo", ThermoDetails),
                ("bcond", BCSettings * 5), ("rhoMin", c_double)]


class ParametricRadiation(ctypes.Structure):
    _fields_ = [("exponent", c_double), ("prefactor", c_double), ("minTemperature", c_double)]


STRUCTS = {"BCSettings": BCSettings, "MGArray": MGArray, "GeometryParams": GeometryParams,
           "ThermoDetails": ThermoDetails, "FluidStepParams": FluidStepParams,
           "ParallelTopology": ParallelTopology, "DragModelNorm": DragModelNorm, "GravityData": GravityData,
           "GridFluid": GridFluid, "ParametricRadiation": ParametricRadiation,
           "ImogenBoundaryData": ImogenBoundaryData}

_P = ctypes.POINTER
_DP = _P(c_double)
_I64P = _P(c_int64)
_IP = _P(c_int)
_MGP = _P(MGArray)
_FP = _P(ctypes.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "calcPartitionExtent": (None, [_MGP, c_int, _IP]),
    "deserializeTagToMGArray": (c_int, [_I64P, _MGP]),
    "serializeMGArrayToTag": (None, [_MGP, _I64P]),
    "MGA_permtagToNums": (None, [c_int, _IP]),
    "MGA_numsToPermtag": (c_int, [_IP]),
    "MGA_dir2memdir": (c_int, [_IP, c_int]),
    "MGA_allocArrays": (c_int, [_P(_MGP), c_int, _MGP]),
    "MGA_allocSlab": (c_int, [_MGP, _MGP, c_int]),
    "MGA_delete": (c_int, [_MGP]),
    "MGA_uploadArrayToGPU": (c_int, [_DP, _MGP, c_int]),
    "MGA_downloadArrayToCPU": (c_int, [_MGP, _P(_DP), c_int]),
    "GPU_upload": (c_int, [_DP, _IP, c_int, c_int, _I64P]),
    "GPU_download": (c_int, [_I64P, _DP]),
    "GPU_free": (c_int, [_I64P]),
    "GPU_makeslab": (c_int, [_I64P, c_int, _I64P]),
    "GPU_getslab": (c_int, [_I64P, c_int, _I64P]),
    "GPU_setslab": (c_int, [_I64P, c_int, _DP, _I64P]),
    "GPU_clone": (c_int, [_I64P, _I64P]),
    "GPU_copy": (c_int, [_I64P, _I64P]),
    "GPU_ctrl_memory": (c_int, [c_int, _P(ctypes.c_size_t), _P(ctypes.c_size_t)]),
    "GPU_ctrl_reset": (c_int, []),
    "GPU_ctrl_info": (c_int, [c_int, ctypes.c_char_p, c_int]),
    "performFluidUpdate_3D": (c_int, [_MGP, _P(ParallelTopology), FluidStepParams, _MGP]),
    "performFluidUpdate_3D_noalloc": (c_int, [_MGP, _P(ParallelTopology), FluidStepParams]),
    "performFluidUpdate_1D": (c_int, [_MGP, FluidStepParams, _P(ParallelTopology), _MGP]),
    "setFluidBoundary": (c_int, [_MGP, _P(GeometryParams), c_int]),
    "setArrayBoundaryConditions": (c_int, [_MGP, _P(GeometryParams), c_int, c_int]),
    "setFreeBalanceCondition": (c_int, [_MGP, _P(GeometryParams), c_int, c_int, _MGP, c_double]),
    "applyStaticCells": (c_int, [_MGP, _MGP, _P(c_double)]),
    "calculateSoundspeed": (c_int, [_MGP, _MGP, _MGP, c_double]),
    "computeLocalCFLTimestep": (c_int, [_MGP, _MGP, _P(GeometryParams), c_int, _DP, _DP]),
    "imogen_cflTimestepFused": (c_int, [_MGP, _P(GeometryParams), c_int, c_double, _DP, _DP]),
    "alterArrayMetadata": (c_int, [_MGP, _MGP, c_int]),
    "flipArrayIndices": (c_int, [_MGP, _P(_MGP), c_int, c_int, c_void_p, _MGP]),
    "exchange_MPI_Halos": (c_int, [_MGP, c_int, _P(ParallelTopology), c_int]),
    "performSourceFunctions": (c_int, [c_int, _P(GridFluid), c_int, FluidStepParams, _P(ParallelTopology),
                                       _P(GravityData), _P(ParametricRadiation), _MGP]),
    "sourcefunction_2FluidDrag": (c_int, [_MGP, _MGP, _P(GeometryParams), _P(ThermoDetails),
                                          _P(ThermoDetails), c_double, c_int]),
    "imogen_drag_set_model": (c_int, [_FP, _P(DragModelNorm)]),
    "imogen_drag_learned_stats": (c_int, [_I64P, _I64P]),
    "imogen_drag_training_samples": (c_int, [_P(ThermoDetails), _P(ThermoDetails), c_int64, c_void_p, c_void_p,
                                             c_void_p]),
    "imogen_mlp_forward": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "imogen_mlp_backward": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "imogen_mlp_adam": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, c_int]),
    "imogen_mlp_infer_bf16": (c_int, [c_int64, c_void_p, c_void_p]),
    "imogen_mlp_mse_grad": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "imogen_gather_rows": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    "imogen_ensemble_create": (c_int, [c_int, _IP, _IP, _DP, _DP, c_double, c_int, c_double, _P(c_void_p)]),
    "imogen_ensemble_set_state": (c_int, [c_void_p, c_int, _DP]),
    "imogen_ensemble_get_state": (c_int, [c_void_p, c_int, _DP]),
    "imogen_ensemble_run": (c_int, [c_void_p, c_int, c_int, c_int]),
    "imogen_ensemble_times": (c_int, [c_void_p, _DP]),
    "imogen_ensemble_autotune": (c_int, [c_void_p, c_int, _IP]),
    "imogen_ensemble_x_lanes": (c_int, [c_void_p, _IP, _IP]),
    "imogen_ensemble_destroy": (c_int, [c_void_p]),
    "imogen_ensemble_set_dust": (c_int, [c_void_p, _IP, c_double, _P(ThermoDetails), _P(ThermoDetails), c_int]),
    "imogen_ensemble_set_dust_state": (c_int, [c_void_p, c_int, _DP]),
    "imogen_ensemble_get_dust_state": (c_int, [c_void_p, c_int, _DP]),
    "imogen_timeloop_create": (c_int, [_MGP, c_int, _P(GeometryParams), _P(FluidStepParams), _DP, c_double,
                                       c_double, _P(c_void_p)]),
    "imogen_timeloop_run": (c_int, [c_void_p, c_int, c_int, c_int]),
    "imogen_timeloop_set_drag": (c_int, [c_void_p, _P(ThermoDetails), _P(ThermoDetails), c_int]),
    "imogen_timeloop_clock": (c_int, [c_void_p, _DP, _DP]),
    "imogen_timeloop_set_time": (c_int, [c_void_p, c_double]),
    "imogen_timeloop_destroy": (c_int, [c_void_p]),
    "imogen_device_alloc": (c_int, [ctypes.c_size_t, _P(c_void_p)]),
    "imogen_device_free": (c_int, [c_void_p]),
    "imogen_memcpy_htod": (c_int, [c_void_p, c_void_p, ctypes.c_size_t]),
    "imogen_memcpy_dtoh": (c_int, [c_void_p, c_void_p, ctypes.c_size_t]),
    "imogen_profile_enable": (c_int, [c_int]),
    "imogen_profile_read": (c_int, [c_int, _I64P, _DP]),
    "imogen_stream": (c_void_p, []),
    "imogen_synchronize": (c_int, []),
    "imogen_check_errors": (c_int, []),
    "imogen_last_error": (ctypes.c_char_p, []),
    "imogen_version": (ctypes.c_char_p, []),
}

_lib = None


class ImogenError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} returned {code}: {msg}")
        self.code = code


def lib():
    """Loads libimogen_hip.so. There is no fallback: a missing library raises."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built; run __graft_entry__.build() "
                              "or make -C gpuimogen_amd/csrc")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn, code):
    if code != SUCCESSFUL:
        raise ImogenError(fn, code, lib().imogen_last_error().decode())
    return code


def call(name, *args):
    return check(name, getattr(lib(), name)(*args))
