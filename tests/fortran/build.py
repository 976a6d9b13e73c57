"""Builds the Fortran host test program tests/fortran/_build/pede_host against libmpgpu.so
through include/mpgpu_mod.f90 (flang; the library is found next to the program by rpath)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")


def flang():
    for c in ("/opt/rocm/llvm/bin/flang", shutil.which("flang") or "", shutil.which("amdflang") or ""):
        if c and os.path.exists(c):
            return c
    return None


def build():
    fc = flang()
    if fc is None:
        raise SystemExit("flang not found: the Fortran host test program cannot be built")
    os.makedirs(OUT, exist_ok=True)
    mod = os.path.join(ROOT, "include", "mpgpu_mod.f90")
    src = os.path.join(HERE, "pede_host.f90")
    lib = os.path.join(ROOT, "millepede-ii_amd")
    exe = os.path.join(OUT, "pede_host")
    cmds = [
        [fc, "-O2", "-J", OUT, "-c", mod, "-o", os.path.join(OUT, "mpgpu_mod.o")],
        [fc, "-O2", "-I", OUT, "-c", src, "-o", os.path.join(OUT, "pede_host.o")],
        [fc, os.path.join(OUT, "pede_host.o"), os.path.join(OUT, "mpgpu_mod.o"), "-L" + lib, "-lmpgpu",
         "-Wl,-rpath,$ORIGIN/../../../millepede-ii_amd", "-o", exe + ".tmp"],
    ]
    for c in cmds:
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            raise SystemExit(" ".join(c) + "\n" + r.stdout + r.stderr)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build())
