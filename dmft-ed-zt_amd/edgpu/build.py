"""Build recipe of libedgpu.so (hipcc, gfx950), kept beside the package.

Usage: python -m edgpu.build   (from dmft-ed-zt_amd/), or build() from Python.
The library is written in-tree next to this file so that it travels with
the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(HERE, "libedgpu.so")
SOURCES = ["kernels.hip", "batch.hip", "noalloc.hip", "reorth.hip", "capi.cpp", "batch.cpp", "tridiag.cpp",
           "devmem.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 (v_pk_*_f32) stays off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"),
]


FORTRAN = os.path.join(PKG, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
EXAMPLE = os.path.join(FORTRAN, "example_solve")
EXAMPLE_BATCH = os.path.join(FORTRAN, "example_batch")


def build_fortran(verbose=False):
    """Compile the iso_c_binding module and the example caller (amdflang)."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(FORTRAN, "edgpu_c.f90")
    cmds = [[FLANG, "-c", mod, "-module-dir", FORTRAN, "-o", os.path.join(FORTRAN, "edgpu_c.o")]]
    for src, exe in (("example_solve.f90", EXAMPLE), ("example_batch.f90", EXAMPLE_BATCH)):
        cmds.append([FLANG, "-I", FORTRAN, os.path.join(FORTRAN, src), os.path.join(FORTRAN, "edgpu_c.o"),
                     "-L", HERE, "-ledgpu", "-Wl,-rpath,$ORIGIN/../edgpu", "-o", exe])
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd, cwd=FORTRAN)
    return EXAMPLE


def build(verbose=False):
    """Compile every HIP/C++ source into one shared library (always rebuilds)."""
    objs, procs = [], []
    for src in SOURCES:  # translation units compile in parallel
        obj = os.path.join(CSRC, src + ".o")
        cmd = [HIPCC, *FLAGS, "-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, pr in procs:
        if pr.wait() != 0:
            raise subprocess.CalledProcessError(pr.returncode, cmd)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
