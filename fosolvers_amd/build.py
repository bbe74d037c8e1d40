"""Build recipe for libfosolvers_amd.so (gfx950) -- hipcc for the device sources, g++ flags for
the host C++ sources, one shared library in the package directory.

    python -m fosolvers_amd.build          # or fosolvers_amd.build.build()
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfosolvers_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["euler_kernels.hip", "context.hip", "precond.hip", "bdf.hip", "learned.hip", "mesh_gpu.hip",
               "ilut.hip", "vtk.hip"]
CPP_SOURCES = ["mesh.cpp", "io.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
HIP_FLAGS = COMMON + [
    f"--offload-arch={ARCH}",
    "-x", "hip",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


FORTRAN = os.path.join(HERE, "fortran")
FOEULER = os.path.join(HERE, "foeuler_amd")


def flang():
    """amdflang (ROCm's Fortran compiler) if present, else None."""
    for c in (os.path.join(ROCM, "bin", "amdflang"), os.path.join(ROCM, "llvm", "bin", "amdflang")):
        if os.path.exists(c):
            return c
    return None


def build_fortran(verbose=False):
    """The Fortran interface module and the foeuler driver (fosolvers_amd/foeuler_amd), linked
    against the library with an rpath to the package directory.  Skipped when amdflang is absent."""
    fc = flang()
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran driver not built")
        return None
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    tmp = FOEULER + ".tmp"
    _run([fc, "-O2", "-module-dir", mod, os.path.join(FORTRAN, "modFosAmd.f90"),
          os.path.join(FORTRAN, "foeuler_amd.f90"), "-o", tmp, "-L" + HERE, "-lfosolvers_amd",
          "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, FOEULER)
    if verbose:
        print("built", FOEULER)
    return FOEULER


def build(verbose=False):
    """Compile every source and link libfosolvers_amd.so (and the Fortran driver); always rebuilds."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = HIP_SOURCES + CPP_SOURCES
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
