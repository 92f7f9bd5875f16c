"""Builds libclawpack_amd.so (HIP, gfx950) in-tree with hipcc."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclawpack_amd.so")
SOURCES = ["claw2d.hip", "adjoint.hip"]
HEADERS = ["physics.hpp", "adjoint.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b + c is fused into an FMA within one source expression (the front end decides), never
    # across statements by the backend: backend fusion depends on basic-block boundaries, and the
    # step kernel's row body exists in two copies with different blocks (steady and edge rows),
    # which must round alike for results not to depend on a cell's position in its segment
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function",
]


def build_library(verbose=False):
    """Compile every HIP source into clawpack_amd/libclawpack_amd.so (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
