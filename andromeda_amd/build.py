"""Builds libandromeda_hip.so (HIP sources for gfx950) in-tree, next to this file."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libandromeda_hip.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["andr_assemble.hip", "andr_lsqr.hip", "andr_gmres.hip", "andr_resistance.hip", "andr_learned.hip", "andr_io.hip",
           "andr_capi.hip", "andr_prof.hip", "andr_noalloc.hip", "andr_laplace.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def hipcc() -> str:
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(p) or p == "hipcc":
            return p
    return "hipcc"


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_HOST = os.path.join(FORTRAN_DIR, "resistance_host")


def amdflang() -> str | None:
    for p in ("/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/amdflang"):
        if os.path.exists(p):
            return p
    return None


def build_fortran(verbose: bool = True) -> str | None:
    """The ISO_C_BINDING module and the minimal Fortran host, linked against the library."""
    fc = amdflang()
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran host not built")
        return None
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    iface = os.path.join(mod, "andromeda_hip_iface.o")
    steps = [[fc, "-O2", "-J", mod, "-c", os.path.join(FORTRAN_DIR, "andromeda_hip_iface.f90"), "-o", iface],
             [fc, "-O2", "-J", mod, os.path.join(FORTRAN_DIR, "resistance_host.f90"), iface, "-L" + HERE,
              "-landromeda_hip", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_HOST]]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{r.stderr[-6000:]}")
    if verbose:
        print(f"built {FORTRAN_HOST}")
    return FORTRAN_HOST


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 6)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
