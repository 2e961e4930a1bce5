"""In-tree build of the device library csrc/libdfn.so for gfx950.

hipcc compiles the HIP translation units (kernels, engine + C ABI, device
H-matrix construction) for
--offload-arch=gfx950; g++ compiles the host-only setup (geometry, crossings,
H-matrix construction with OpenMP).  Floating-point contraction is off in the
host setup so the operator entries round as the reference's expressions do.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(CSRC, "libdfn.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")
FORT = os.path.join(HERE, "fortran")
CXX = os.environ.get("CXX_HOST", "g++")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=fast-honor-pragmas"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-mavx2",
              "-Wall", "-Wno-unused-function"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)


def build(verbose=True):
    objs = []
    for src in ("host_setup.cpp", "h5min.cpp"):
        o = os.path.join(CSRC, src.replace(".cpp", ".o"))
        _run([CXX] + HOST_FLAGS + ["-c", src, "-o", o])
        objs.append(o)
    for src in ("kernels.hip", "engine.hip", "hbuild.hip", "ensemble.hip", "mlp.hip"):
        o = os.path.join(CSRC, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_fortran()
    return LIB


def build_fortran():
    """The ISO_C_BINDING layer (fortran/dfn_hip.f90) and the Fortran host drivers
    that call the library through it (fortran/drive_solver.f90: one problem;
    fortran/drive_ensemble.f90: an ensemble with the proposal network)."""
    bld = os.path.join(FORT, "build")
    os.makedirs(bld, exist_ok=True)
    mod_o = os.path.join(bld, "dfn_hip.o")
    subprocess.run([FLANG, "-O2", "-J", bld, "-c", os.path.join(FORT, "dfn_hip.f90"), "-o", mod_o], check=True)
    for prog in ("drive_solver", "drive_ensemble"):
        subprocess.run([FLANG, "-O2", "-J", bld, "-I", bld, os.path.join(FORT, prog + ".f90"), mod_o,
                        "-L" + CSRC, "-ldfn", "-Wl,-rpath,$ORIGIN/../../csrc", "-o", os.path.join(bld, prog)],
                       check=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
