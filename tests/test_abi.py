"""The C-ABI library builds for gfx950, loads on a host without a GPU and
exports every entry point include/pnfam_gpu.h declares."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pnfam_gpu.h")
LIB = os.path.join(ROOT, "pynfam_amd", "libpnfam_gpu.so")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pnfam_gpu_\w+)\s*\(", txt)))


def test_library_exports_header_symbols():
    assert os.path.exists(LIB), "build() must have produced the library"
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s


def test_version_and_error_string_without_device():
    lib = ctypes.CDLL(LIB)
    lib.pnfam_gpu_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.pnfam_gpu_version()
    lib.pnfam_gpu_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.pnfam_gpu_create(None, None) == -1            # PNFAM_GPU_EINVAL, no device touched
    lib.pnfam_gpu_last_error.restype = ctypes.c_char_p
    assert lib.pnfam_gpu_last_error()


def test_engine_refuses_to_run_without_library(tmp_path, monkeypatch):
    import pynfam_amd.engine as eng
    monkeypatch.setattr(eng, "_lib", None)
    import pytest
    with pytest.raises(eng.GpuError):
        eng.load_library(str(tmp_path / "missing.so"))


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="no Fortran compiler")
def test_fortran_binding_compiles_and_links(tmp_path):
    """include/pnfam_gpu.f90 (ISO_C_BINDING interfaces) compiles, links against
    the library and calls it without a device (version, argument check)."""
    inc = os.path.join(ROOT, "include", "pnfam_gpu.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "check_binding.f90")
    libdir = os.path.join(ROOT, "pynfam_amd")
    subprocess.check_call(["amdflang", "-c", inc, "-o", "pnfam_gpu.o"], cwd=tmp_path)
    subprocess.check_call(["amdflang", prog, "pnfam_gpu.o", "-L" + libdir, "-lpnfam_gpu", "-Wl,-rpath," + libdir,
                           "-o", "chk"], cwd=tmp_path)
    out = subprocess.run([str(tmp_path / "chk")], capture_output=True, text=True, check=True).stdout
    assert "version: pnfam_gpu" in out and "gfx950" in out
    assert "create rc: -1" in out
    assert "nb must be positive" in out
