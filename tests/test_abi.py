"""CPU checks of the C ABI: the library loads, exports every function declared
in include/imogen_abi.h, the ctypes mirror matches the header's struct layout
(compiled probe), and tag serialisation round-trips (no device calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from gpuimogen_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(abi.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*\b([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "for", "while", "defined")))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signature table covers them all
    assert set(names) <= set(abi.SIGNATURES), set(names) - set(abi.SIGNATURES)


def test_struct_layout_matches_header(tmp_path):
    fields = {name: [f[0] for f in cls._fields_] for name, cls in abi.STRUCTS.items()}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{abi.HEADER}"', "int main(void){"]
    for name, fl in fields.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f in fl:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for name, cls in abi.STRUCTS.items():
        assert int(got[f"{name} size"]) == ctypes.sizeof(cls), name
        for f in fields[name]:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, f"{name}.{f}"


def test_tag_round_trip_and_partition_extent():
    L = abi.lib()
    tag = np.zeros(abi.TAG_LEN_1GPU, dtype=np.int64)
    tag[:11] = [40, 30, 20, 5, 4, 1, 1, 0, 1, 0, 2]
    tag[11:] = [0, 0x7F0000001000]
    m = abi.MGArray()
    assert L.deserializeTagToMGArray(tag.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(m)) == 0
    assert tuple(m.dim) == (40, 30, 20) and m.numel == 24000 and m.numSlabs == 5
    assert m.slabPitch[0] == ((24000 * 8 + 255) // 256) * 256
    assert m.vectorComponent == 2 and tuple(m.currentPermutation) == (1, 2, 3)
    assert m.devicePtr[0] == 0x7F0000001000
    back = np.zeros_like(tag)
    L.serializeMGArrayToTag(ctypes.byref(m), back.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    assert np.array_equal(back, tag)
    sub = (ctypes.c_int * 6)()
    L.calcPartitionExtent(ctypes.byref(m), 0, sub)
    assert tuple(sub) == (0, 0, 0, 40, 30, 20)


def test_permtag_maps_are_inverse():
    L = abi.lib()
    for t in range(1, 7):
        p = (ctypes.c_int * 3)()
        L.MGA_permtagToNums(t, p)
        assert sorted(p) == [1, 2, 3]
        assert L.MGA_numsToPermtag(p) == t
        for d in (1, 2, 3):
            assert p[L.MGA_dir2memdir(p, d) - 1] == d


@pytest.mark.parametrize("bad", [[0, 1, 1], [4, -1, 1]])
def test_malformed_tag_is_rejected(bad):
    L = abi.lib()
    tag = np.zeros(abi.TAG_LEN_1GPU, dtype=np.int64)
    tag[:3] = bad
    tag[6] = 1
    m = abi.MGArray()
    rc = L.deserializeTagToMGArray(tag.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(m))
    assert rc == abi.ERROR_DESERIALIZE_GPUTAG_FAILED


def test_too_many_gpus_rejected():
    L = abi.lib()
    tag = np.zeros(abi.GPU_TAG_LENGTH + 2 * 5, dtype=np.int64)
    tag[:3] = [4, 4, 4]
    tag[6] = 5
    m = abi.MGArray()
    rc = L.deserializeTagToMGArray(tag.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(m))
    assert rc == abi.ERROR_DESERIALIZE_GPUTAG_FAILED


def test_source_step_dispatch_errors_without_device_work():
    """performSourceFunctions (sourceStep.cu:200-226) rejects what is not
    built before touching any array: types with cylindrical, rotating-frame or
    gravity terms, radiation, unknown types, a one-fluid drag and a missing
    radiation record. Type 0 without radiation does nothing."""
    L = abi.lib()
    fsp = abi.FluidStepParams()
    rad = abi.ParametricRadiation()
    gf = (abi.GridFluid * 2)()
    P = ctypes.byref

    def call(src, n=2, r=rad):
        return L.performSourceFunctions(src, gf, n, fsp, None, None, P(r) if r is not None else None, None)

    assert call(0) == 0
    for src in (1, 2, 3, 4, 5, 6, 7, 9, 12, 15):
        assert call(src) == abi.ERROR_NOIMPLEMENT, src
    assert call(16) == abi.ERROR_INVALID_ARGS and call(-1) == abi.ERROR_INVALID_ARGS
    assert call(8, n=1) == abi.ERROR_INVALID_ARGS
    assert call(8, r=None) == abi.ERROR_NULL_POINTER
    rad.prefactor = 1.0
    assert call(0) == abi.ERROR_NOIMPLEMENT and call(8) == abi.ERROR_NOIMPLEMENT


def test_index_exchange_metadata_follows_flux_m_tables():
    """alterArrayMetadata (cudaArrayRotateB.cu:583-640) walked over flux.m's own
    preperm / permcall tables (flux/flux.m:29-40): before every sweep the
    swept direction fluxcall(n) sits in memory x, and every sequence ends in
    XYZ order (permtag 1) with the dims restored; unknown codes are rejected."""
    from gpuimogen_amd import gpuclass as gc
    L = abi.lib()
    for s in range(6):
        m = abi.MGArray()
        m.dim[0], m.dim[1], m.dim[2] = 16, 12, 14
        m.nGPUs = 1
        m.currentPermutation[0], m.currentPermutation[1], m.currentPermutation[2] = 1, 2, 3
        m.permtag = 1
        m.partitionDir = abi.PARTITION_X
        m.partNumel[0] = 16 * 12 * 14
        if gc._PREPERM[s]:
            assert L.alterArrayMetadata(ctypes.byref(m), ctypes.byref(m), gc._PREPERM[s]) == 0
        for n in range(3):
            d = gc._FLUXCALL[n][s]
            assert m.currentPermutation[0] == d, (s, n)
            assert m.dim[0] == (16, 12, 14)[d - 1]
            assert m.partNumel[0] == 16 * 12 * 14
            if gc._PERMCALL[n][s]:
                assert L.alterArrayMetadata(ctypes.byref(m), ctypes.byref(m), gc._PERMCALL[n][s]) == 0
        assert tuple(m.currentPermutation) == (1, 2, 3) and m.permtag == 1
        assert tuple(m.dim) == (16, 12, 14) and m.partitionDir == abi.PARTITION_X
    m = abi.MGArray()
    assert L.alterArrayMetadata(ctypes.byref(m), ctypes.byref(m), 7) == abi.ERROR_INVALID_ARGS
