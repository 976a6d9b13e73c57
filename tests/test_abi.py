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


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang not available")
def test_fortran_program_calls_library():
    """A Fortran caller (as pede would be) links libmpgpu.so through the ISO_C_BINDING module and
    reads pede's defaults back (no device call)."""
    src = os.path.join(ROOT, "include", "mpgpu_mod.f90")
    prog = """
PROGRAM t
    USE mpgpu_mod
    IMPLICIT NONE
    TYPE(mpgpu_options) :: o
    CALL mpgpu_default_options(o)
    PRINT '(3I8,F8.2)', o%itnlim, o%mreqenf, o%lfitbb, o%chhuge
END PROGRAM t
"""
    libdir = os.path.dirname(mpgpu.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "t.f90"), "w") as f:
            f.write(prog)
        r = subprocess.run(["flang", "-J", d, src, os.path.join(d, "t.f90"), "-L" + libdir, "-lmpgpu",
                            "-Wl,-rpath," + libdir, "-o", os.path.join(d, "t")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([os.path.join(d, "t")], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        assert r.stdout.split() == ["2000", "25", "1", "50.00"]


def test_steering_keyword_exit_codes(tmp_path):
    """The steering parser returns pede's exit codes before any device work: 13 for a keyword
    pede does not know (src/pede.f90:10886-10896, 10015-10020), 40 for a pede keyword that would
    change the result and is not part of this path, 15 for a missing binary file."""
    import mpgpu
    cases = {"unknown": ("notakeyword 3\nend\n", 13), "unsupported": ("weightedcons\nend\n", 40),
             "pairs": ("pairentries 5\nend\n", 40), "missing": ("nosuchfile.bin\nend\n", 15)}
    for name, (text, code) in cases.items():
        f = tmp_path / f"{name}.txt"
        f.write_text(text)
        with pytest.raises(mpgpu.MpgpuError) as e:
            mpgpu.run_steering(str(f), str(tmp_path))
        assert e.value.code == code, (name, e.value)
        assert (tmp_path / "millepede.end").read_text().split()[0] == str(code)
