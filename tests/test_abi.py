"""The C-ABI library loads and exports every entry point declared in include/mpgpu.h;
the Fortran binding module compiles.  CPU only (no compute calls)."""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import pytest

import mpgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "mpgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpgpu_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_wrapper():
    assert declared() == sorted(mpgpu.EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(mpgpu.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_generator_library_loads():
    g = mpgpu.gen()
    for name in ("mpgen_toy", "mpgen_pixel", "mpgen_write", "mpgen_free"):
        assert hasattr(g, name)


def test_default_options_are_pedes():
    o = mpgpu.default_options()
    assert (o.mreqenf, o.mreqena, o.itnlim, o.lfitbb) == (25, 10, 2000, 1)
    assert abs(o.chhuge - 50.0) < 1e-6 and abs(o.cndlmx - 100.0) < 1e-6 and abs(o.mrestl - 1e-6) < 1e-18


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang not available")
def test_fortran_binding_compiles():
    src = os.path.join(ROOT, "include", "mpgpu_mod.f90")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run(["flang", "-c", src, "-J", d, "-o", os.path.join(d, "m.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
