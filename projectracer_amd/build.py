"""Build libprojectracer_gpu.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m projectracer_amd.build   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libprojectracer_gpu.so"
SOURCES = ["pt_setup.cpp", "pt_solver.hip", "pt_direct.hip", "pt_tracers.hip", "pt_equalize.hip", "pt_sort.hip",
           "pt_stats.hip", "pt_files.cpp", "pt_api.hip", "pt_unet.hip", "pt_ensemble.cpp"]

# gfx950 only.  -ffp-contract=off keeps every FP64 expression in the
# reference's operation order (no fused multiply-add contraction); packed
# FP32 instructions are disabled for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        s = CSRC / src
        o = objdir / (s.stem + ".o")
        lang = ["-x", "hip"] if s.suffix == ".hip" else []
        cmd = [hipcc(), *HIPCC_FLAGS, *lang, "-c", str(s), "-o", str(o)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        if r.stderr.strip() and verbose:
            print(r.stderr, file=sys.stderr)
        objs.append(str(o))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = PKG / "fortran"
FORTRAN_PROG = LIBDIR / "pt_steps"


def fortran_compiler() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/llvm/bin/flang", shutil.which("flang"), shutil.which("gfortran")):
        if cand and Path(cand).exists():
            return cand
    return None


def build_fortran(verbose: bool = False) -> Path | None:
    """The Fortran time-loop program (fortran/pt_steps.f90) linked against
    the library: the interface module used as a reference maintainer would.
    Returns None when the image has no Fortran compiler."""
    fc = fortran_compiler()
    if fc is None:
        return None
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    mod_o = objdir / "projectracer_gpu_f90.o"
    tmp = FORTRAN_PROG.with_name(FORTRAN_PROG.name + ".tmp")
    for cmd in ([fc, "-O2", "-J", str(objdir), "-c", str(FORTRAN_DIR / "projectracer_gpu.f90"), "-o", str(mod_o)],
                [fc, "-O2", "-J", str(objdir), str(FORTRAN_DIR / "pt_steps.f90"), str(mod_o), f"-L{LIBDIR}",
                 "-lprojectracer_gpu", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, FORTRAN_PROG)
    return FORTRAN_PROG


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
