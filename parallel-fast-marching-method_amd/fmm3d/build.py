"""Build the native parts in-tree: libfmm3d_hip.so (hipcc, gfx950) and the
drop-in Fortran host fm3dt_gpu (flang).  Built files stay next to the
sources so they travel with the repository snapshot."""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                 # parallel-fast-marching-method_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
HOST = os.path.join(ROOT, "host")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libfmm3d_hip.so")
EXE = os.path.join(ROOT, "host", "fm3dt_gpu")
INVERT_EXE = os.path.join(ROOT, "host", "fm3dt_invert")
SLICE_EXE = os.path.join(ROOT, "host", "gmtslice_gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # float64 results must not depend on FMA contraction (they are compared bit for bit)
    "-ffp-contract=off",
    # no packed float32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

SOURCES = ["fmm3d_solver.hip", "fmm3d_server.hip", "fmm3d_slice.hip", "fmm3d_gmat.hip"]


def build_hip(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_host() -> str:
    """Fortran hosts (ISO_C_BINDING onto libfmm3d_hip.so): the drop-in
    fm3dt_gpu, the resident inversion loop fm3dt_invert and the model
    slicer gmtslice_gpu, all on the shared module file fm3dt_core.f90."""
    core = os.path.join(HOST, "fm3dt_core.f90")
    if not os.path.exists(core) or not os.path.exists(FLANG):
        return ""
    moddir = os.path.join(HOST, "_mod")
    os.makedirs(moddir, exist_ok=True)
    for main, exe in (("fm3dt_gpu.f90", EXE), ("fm3dt_invert.f90", INVERT_EXE), ("gmtslice_gpu.f90", SLICE_EXE)):
        cmd = [FLANG, "-O2", "-J", moddir, core, os.path.join(HOST, main), "-o", exe, "-L" + PKG, "-lfmm3d_hip",
               "-Wl,-rpath,$ORIGIN/../fmm3d", "-Wl,-rpath," + PKG]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    return EXE


def build_all() -> None:
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
    print(LIB)
