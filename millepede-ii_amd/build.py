"""Build libmpgpu.so (HIP, gfx950) and libmpgen.so (host C++ generators) in place.

Usage: python millepede-ii_amd/build.py [--jobs N]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("MPGPU_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # keep packed-FP32 VALU instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp", "-Wno-unused-result",
]

HIP_SOURCES = ["k_records.hip", "k_records_band.hip", "k_solve.hip", "k_chol.hip", "k_qlelim.hip", "k_struct.hip", "mpgpu.cpp",
               "mpgpu_qldec.cpp", "mpgpu_pede.cpp", "mpgpu_serve.cpp"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit(f"build failed: {cmd[-1]}")
    return r


def compile_one(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def build(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    lib = os.path.join(HERE, "libmpgpu.so")
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp", "-o", lib + ".tmp"] + objs + ["-lz"])
    os.replace(lib + ".tmp", lib)
    gen = os.path.join(HERE, "libmpgen.so")
    run(["g++", "-O3", "-fPIC", "-shared", "-std=c++17", "-fopenmp", "-o", gen + ".tmp",
         os.path.join(CSRC, "mpgen.cpp")])
    os.replace(gen + ".tmp", gen)
    return lib, gen


if __name__ == "__main__":
    j = 8
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    print(build(j))
