"""Build libx3d.so (the HIP/gfx950 C-ABI library) in-tree with hipcc.

``python -m incompact3d_amd.build`` or ``incompact3d_amd.build.build()``.
Objects go to ``incompact3d_amd/_build/``; the library to
``incompact3d_amd/libx3d.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libx3d.so")
SOURCES = ["capi.cpp", "schemes.cpp", "lines_f64.hip", "lines_f32.hip", "fft.hip", "fields.hip", "solver.hip",
           "momentum.hip", "projection.hip", "les.hip", "mlp.hip", "capi_les.cpp", "channel.hip", "capi_host.cpp"]
ARCH = os.environ.get("X3D_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops keeps the compiler from emitting packed FP32 arithmetic
# (v_pk_add/mul/fma_f32); -ffp-contract=fast lets
# hipcc fuse multiply-adds.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-ffp-contract=fast", "-Wall", "-Wno-unused-function",
            "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return obj


FORTRAN = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FORTRAN, "_build")
FCHECK = os.path.join(FBUILD, "check_binding")


def flang() -> str | None:
    for cand in (os.environ.get("X3D_FLANG"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"),
                 shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran binding (x3d_binding, x3d_drop_in) and the
    check_binding caller against libx3d.so; None when no Fortran compiler is
    installed."""
    fc = flang()
    if fc is None:
        if verbose:
            print("no Fortran compiler: the Fortran binding is not built", file=sys.stderr)
        return None
    os.makedirs(FBUILD, exist_ok=True)
    steps = [[fc, "-O2", "-J", FBUILD, "-c", os.path.join(FORTRAN, "x3d_binding.f90"), "-o",
              os.path.join(FBUILD, "x3d_binding.o")],
             [fc, "-O2", "-J", FBUILD, "-c", os.path.join(FORTRAN, "x3d_drop_in.F90"), "-o",
              os.path.join(FBUILD, "x3d_drop_in.o")],
             [fc, "-O2", "-J", FBUILD, "-c", os.path.join(FORTRAN, "check_binding.F90"), "-o",
              os.path.join(FBUILD, "check_binding.o")],
             [fc, os.path.join(FBUILD, "check_binding.o"), os.path.join(FBUILD, "x3d_drop_in.o"),
              os.path.join(FBUILD, "x3d_binding.o"), "-L" + HERE, "-lx3d", "-Wl,-rpath,$ORIGIN/../..", "-o",
              FCHECK]]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return FCHECK


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
