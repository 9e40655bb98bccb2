"""gpuimogen_amd -- MI355X (gfx950) implementation of the Imogen fluid step.

The product is the C-ABI library libimogen_hip.so (include/imogen_abi.h);
this package is the host-side mirror of the reference's MATLAB call forms
(gpuclass.py) and the ctypes binding (abi.py). There is no CPU fallback: if
the library is missing or the GPU is absent, calls raise.
"""
from . import abi, gpuclass, problems  # noqa: F401

__all__ = ["abi", "gpuclass", "problems"]
