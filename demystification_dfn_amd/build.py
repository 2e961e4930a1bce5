"""In-tree build of the device library csrc/libdfn.so for gfx950.

hipcc compiles the HIP translation units (kernels, engine + C ABI) for
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
CXX = os.environ.get("CXX_HOST", "g++")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=fast"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-mavx2",
              "-Wall", "-Wno-unused-function"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)


def build(verbose=True):
    objs = []
    o = os.path.join(CSRC, "host_setup.o")
    _run([CXX] + HOST_FLAGS + ["-c", "host_setup.cpp", "-o", o])
    objs.append(o)
    for src in ("kernels.hip", "engine.hip"):
        o = os.path.join(CSRC, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
