"""Builds libclawpack_amd.so (HIP, gfx950) in-tree with hipcc.

Each source is compiled to an object of its own (in parallel, in a temporary folder), then the
objects are linked into one shared library.  Every object carries its own device code and
registers its own kernels; claw2d.hip launches the kernels of the kernel units through their
handles."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclawpack_amd.so")
HEADERS = ["physics.hpp", "adjoint.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b + c is fused into an FMA within one source expression (the front end decides), never
    # across statements by the backend: backend fusion depends on basic-block boundaries, and the
    # step kernel's row body exists in two copies with different blocks (steady and edge rows),
    # which must round alike for results not to depend on a cell's position in its segment
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function",
]

# (source, extra flags).  The kernel units hold one step kernel each, built with another LLVM
# machine scheduler; it changes instruction order and register assignment, not arithmetic.
SOURCES = [
    ("claw2d.hip", []),
    ("adjoint.hip", []),
    # FP32 shallow water (MC, transverse_waves 2): -2.7% on MI355X
    ("claw2d_swe32.hip", ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]),
    # FP64 Euler (MC, transverse_waves 2): -2.6%, no spill
    ("claw2d_euler64.hip", ["-mllvm", "-amdgpu-sched-strategy=iterative-minreg"]),
    # FP32 Euler (MC, transverse_waves 2): -1.1%
    ("claw2d_euler32.hip", ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]),
]


def build_library(verbose=False):
    """Compile every HIP source into clawpack_amd/libclawpack_amd.so (always recompiles)."""
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="clawpack_amd_build_") as bdir:
        objs, procs = [], []
        for src, extra in SOURCES:
            obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
            cmd = [HIPCC] + FLAGS + extra + ["-c", "-o", obj, os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd)))
            objs.append(obj)
        failed = [src for src, p in procs if p.wait() != 0]
        if failed:
            raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
