"""In-tree build of the MI355X library (libibgpu.so) and the Fortran driver.

Run as ``python -m ib_parallel_amd.build`` or through __graft_entry__.build().
hipcc cross-compiles for gfx950 without a GPU present.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libibgpu.so")
FORTRAN = os.path.join(PKG, "fortran")
DRIVER = os.path.join(PKG, "bin", "ib")

SOURCES = ["ib_fft.hip", "ib_dst_dense.hip", "ib_fluid.hip", "ib_body.hip", "ib_beam.hip", "ib_gmres.hip", "ib_pressure.hip", "ib_subdomain.hip",
           "ib_adjoint.hip", "ib_adjoint_host.hip", "ib_solver.hip", "ib_batch.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _newer(target, deps):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build_lib(force=True, verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    deps = srcs + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(ROOT, "include", "ib_gpu.h"))
    if not force and _newer(LIB, deps):
        return LIB
    tmp = LIB + ".tmp"
    # rocSOLVER (with rocBLAS): the adjoint's LU of the dense Schur matrix;
    # roctx: the host-side ranges of the step's phases
    cmd = [HIPCC] + HIPFLAGS + ["-I", os.path.join(ROOT, "include")] + srcs + ["-lrocsolver", "-lrocblas", "-lrocprofiler-sdk-roctx", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_driver(force=True, verbose=False):
    """PETSc-free Fortran driver (amdflang) linked against libibgpu.so."""
    fc = os.environ.get("FC", shutil.which("amdflang") or "/opt/rocm/bin/amdflang")
    srcs = [os.path.join(FORTRAN, "ib_gpu_mod.f90"), os.path.join(FORTRAN, "ib_main.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    if not force and _newer(DRIVER, srcs + [LIB]):
        return DRIVER
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    moddir = os.path.join(PKG, "bin", "mod")
    os.makedirs(moddir, exist_ok=True)
    cmd = [fc, "-O2", "-module-dir", moddir] + srcs + ["-L", PKG, "-libgpu", "-Wl,-rpath,$ORIGIN/..",
                                                        "-o", DRIVER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return DRIVER


def build(force=True, verbose=False):
    build_lib(force, verbose)
    build_driver(force, verbose)


if __name__ == "__main__":
    build(force="--lazy" not in sys.argv, verbose=True)
