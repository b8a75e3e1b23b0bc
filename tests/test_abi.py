"""The C-ABI libraries load and export every entry point include/abcmb.h declares."""
import ctypes as C
import os
import re
import subprocess

import pytest

from abcmb_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "abcmb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(abcmb_\w+)\s*\(", src))


def _exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_header_entry_points_are_exported():
    decl = _declared()
    host = {d for d in decl if d.startswith("abcmb_host_")}
    gpu = decl - host
    assert gpu and host
    assert gpu <= _exported(abi.HIP_LIB_PATH), gpu - _exported(abi.HIP_LIB_PATH)
    assert host <= _exported(abi.HOST_LIB_PATH)
    assert set(abi.HIP_SYMBOLS) <= gpu


def test_libraries_load_without_gpu():
    lib = C.CDLL(abi.HIP_LIB_PATH)
    for name in abi.HIP_SYMBOLS:
        assert hasattr(lib, name)
    abi.host_lib()


def test_struct_layouts_match_header():
    # the ctypes mirrors must match the C sizes used by the kernels
    assert C.sizeof(abi.AbcmbCosmo) == 400
    assert abi.AbcmbCosmo.fluid_par.offset == 31 * 8 and abi.AbcmbCosmo.pt_fluid.offset == 392
    assert C.sizeof(abi.AbcmbPertItem) == 48
    assert C.sizeof(abi.AbcmbOdeSettings) == 9 * 8 + 10 * 4
    assert C.sizeof(abi.AbcmbTangent) == 48


def test_hot_path_refuses_missing_library(monkeypatch, tmp_path):
    monkeypatch.setattr(abi, "HIP_LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(abi, "_hip", None)
    with pytest.raises(abi.HipLibraryMissing):
        abi.hip_lib()
