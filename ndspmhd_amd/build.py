"""Builds the HIP library ndspmhd_amd/libndsp_hip.so for gfx950 (in-tree).

    python -m ndspmhd_amd.build

hipcc cross-compiles without a GPU.  Packed FP32 instructions are disabled
(-packed-fp32-ops) for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libndsp_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NDSP_OFFLOAD_ARCH", "gfx950")

SOURCES = [os.path.join(CSRC, "ndsp.hip"), os.path.join(CSRC, "ndsp_mlp.hip")]
HEADERS = [os.path.join(CSRC, f) for f in ("ndsp_dev.h", "ndsp_kernels.h", "ndsp_mlp.h")] + [
    os.path.join(ROOT, "include", "ndsp.h")]


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC,
            "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-value",
            "-Wl,-soname,libndsp_hip.so"]


def build(force: bool = True, verbose: bool = False) -> str:
    if not force and os.path.exists(LIB):
        newest = max(os.path.getmtime(p) for p in SOURCES + HEADERS)
        if os.path.getmtime(LIB) >= newest:
            return LIB
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + flags() + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + p.stdout[-8000:] + p.stderr[-8000:])
    os.replace(tmp, LIB)
    return LIB


FDIR = os.path.join(HERE, "fortran")
FSOURCES = [os.path.join(FDIR, f) for f in ("ndsp_capi.f90", "ndsp_host.f90", "ndsp_run.f90")]
FEXE = os.path.join(FDIR, "ndsp_run")
FC = os.environ.get("NDSP_FC", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = False) -> str:
    """Builds the Fortran host program ndspmhd_amd/fortran/ndsp_run against
    libndsp_hip.so (iso_c_binding, no compiler-specific extensions)."""
    import tempfile
    if not os.path.exists(FC):
        raise RuntimeError("Fortran compiler %s not found" % FC)
    with tempfile.TemporaryDirectory() as d:
        cmd = [FC, "-O2", "-module-dir", d] + FSOURCES + [
            LIB, "-Wl,-rpath,$ORIGIN/..", "-o", FEXE + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        p = subprocess.run(cmd, capture_output=True, text=True, cwd=d)
        if p.returncode != 0:
            raise RuntimeError("Fortran host build failed:\n" + p.stdout[-4000:] + p.stderr[-4000:])
    os.replace(FEXE + ".tmp", FEXE)
    return FEXE


if __name__ == "__main__":
    print(build(force=True, verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
