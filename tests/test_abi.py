"""The C-ABI library builds for gfx950, loads on a host without a GPU and
exports every entry point include/pnfam_gpu.h declares."""
import ctypes
import os
import re

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
