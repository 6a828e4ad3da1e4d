"""Build steps for the native parts (called by __graft_entry__.build()).

  libfsim_hip.so   hipcc --offload-arch=gfx950 from fluidsim_amd/csrc/*.hip
  fsim_main        amdflang Fortran host (fluidsim_amd/fortran) linked to it
All outputs are written in-tree so that they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfsim_hip.so")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "_build")
FEXE = os.path.join(FBUILD, "fsim_main")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")

# -ffp-contract=off keeps the reference's rounding sequence (no FMA contraction);
# packed FP32 instructions are disabled (they are unreliable on this pool).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_SOURCES = ["fsim_api.hip", "fsim_kernels.hip", "fsim_learn.hip"]


def build_lib(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_SOURCES = ["fsim_capi.F90", "dev_mirror.F90", "props.F90", "circ_lib.F90",
                   "sfc_lib.F90", "ptcl_lib.F90", "lapl_lib.F90", "main.F90"]


def build_fortran(verbose: bool = False) -> str | None:
    if not os.path.exists(os.path.join(FDIR, "main.F90")):
        return None
    os.makedirs(FBUILD, exist_ok=True)
    objs = []
    for s in FORTRAN_SOURCES:
        src = os.path.join(FDIR, s)
        obj = os.path.join(FBUILD, s.replace(".F90", ".o"))
        cmd = [FC, "-O2", "-c", src, "-o", obj, "-J", FBUILD]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True, cwd=FBUILD)
        objs.append(obj)
    cmd = [FC, "-o", FEXE, *objs, "-L" + HERE, "-lfsim_hip", "-Wl,-rpath," + HERE]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True, cwd=FBUILD)
    return FEXE


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], check=True)


def build_all(verbose: bool = False) -> None:
    build_lib(verbose)
    build_oracle()
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
