"""Builds the HIP engine (libmxg.so) in-tree for gfx950.

The library holds the HIP kernels, the host engine and the C ABI declared in
include/mxg.h.  It is compiled with hipcc directly: no JIT cache, so the
built file travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmxg.so")

SOURCES = ["mx_potential.cpp", "mxg_engine.hip", "mxg_adjoint.hip", "mxg_capi.cpp"]

# -ffp-contract=off: the reference never fuses multiply-adds, and bit-exact
# cell assignment depends on the same rounding.  Packed FP32 instructions
# are disabled for the device code.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False, out=None):
    """Compiles libmxg.so (or the library at out); raises RuntimeError with
    the compiler output."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    out = out or LIB
    tmp = out + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", CSRC] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


def build_driver(src, exe):
    """Compiles a C++ program that drives the engine through include/mxg.hpp
    (host code only, g++), linked against libmxg.so with a run path
    relative to the executable."""
    root = os.path.dirname(HERE)
    rel = os.path.relpath(HERE, os.path.dirname(os.path.abspath(exe)))
    tmp = exe + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src, "-L", HERE, "-lmxg",
           "-Wl,-rpath,$ORIGIN/" + rel, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, exe)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
