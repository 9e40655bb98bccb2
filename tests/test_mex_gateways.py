"""The MEX gateways (mex/*.cpp) keep the reference's MATLAB entry points:
one gateway per MEX name the callers use (imogen.m, relaxingFluid.m,
TimeManager.m, src0001.m, FluidManager / ImogenArray / GPU_Type), each
compiled here against the documented MATLAB C API declarations
(tests/mex_decl/mex.h) and include/imogen_abi.h, and each calling only
entry points that header declares. MATLAB itself is not in this image, so
the gateways are compile-checked, not run."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MEX = os.path.join(ROOT, "mex")
GATEWAYS = ["fluidstep", "cudaFluidStep", "cflTimestep", "cudaSoundspeed", "cudaStatics", "cudaSource2FluidDrag",
            "sourceStep", "GPU_upload", "GPU_download", "GPU_free", "GPU_makeslab", "GPU_getslab", "GPU_setslab",
            "GPU_clone", "GPU_copy", "GPU_ctrl", "cudaArrayRotateB", "cudaHaloExchange"]


def test_every_mex_name_has_a_gateway_and_the_makefile_builds_it():
    mk = open(os.path.join(MEX, "Makefile")).read()
    for g in GATEWAYS:
        assert os.path.exists(os.path.join(MEX, g + ".cpp")), g
        assert re.search(r"\b%s\b" % g, mk), g


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("name", GATEWAYS + ["imogen_mex"])
def test_gateway_compiles_against_the_abi(name):
    src = os.path.join(MEX, name + ".cpp")
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                          "-I" + os.path.join(ROOT, "tests", "mex_decl"), "-I" + os.path.join(ROOT, "include"),
                          "-I" + MEX, src], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_gateways_call_only_declared_entry_points():
    header = open(os.path.join(ROOT, "include", "imogen_abi.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*|double)\s+\*?(\w+)\(", header, re.M))
    helpers = set(re.findall(r"^\s*(?:int|void|double|mxArray \*)\s*\*?(imogen_mex_\w+)\(",
                             open(os.path.join(MEX, "imogen_mex.h")).read(), re.M))
    mx_api = set(re.findall(r"\b(mx\w+|mex\w+)\(", open(os.path.join(ROOT, "tests", "mex_decl", "mex.h")).read()))
    for g in GATEWAYS + ["imogen_mex"]:
        text = open(os.path.join(MEX, g + ".cpp")).read()
        text = re.sub(r"//.*", "", text)
        called = set(re.findall(r"\b([A-Za-z_]\w*)\(", text))
        lib = {c for c in called if c in declared}
        assert lib or g == "imogen_mex" or g == "GPU_ctrl", (g, "calls no library entry point")
        unknown = {c for c in called if (c.startswith("mx") or c.startswith("mex")) and c != "mexFunction"} - mx_api
        assert not unknown, (g, unknown)
        assert all(c in helpers for c in called if c.startswith("imogen_mex_")), g
