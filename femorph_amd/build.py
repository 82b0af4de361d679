"""Build recipe of the femx library: hipcc for gfx950 -> femorph_amd/libfemx.so.

Usage: ``python -m femorph_amd.build`` (also called by __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfemx.so")

SOURCES = ["femx_core.hip", "femx_vms.hip", "femx_supg.hip", "femx_server.hip", "femx_ffd.hip", "femx_ins.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler away from packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-pthread",  # the solve server's prep and GPU threads
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


HOST = os.path.join(CSRC, "host")
MODEL_LIB = os.path.join(HERE, "libfemx_model.so")
DRIVER = os.path.join(HERE, "femx_explicit")
HOST_SOURCES = ["femModel.cpp", "femExplicitFluidSolvers.cpp", "femPointGrid.cpp", "femIncompressibleSolver.cpp"]
GXX_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wno-unused-result"]


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build of {what} failed")


def build_host(verbose: bool = False) -> None:
    """The drop-in C++ model and solver classes (libfemx_model.so, linked to
    libfemx.so) and the femx_explicit driver, with the host compiler."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", HOST]
    rpath = "-Wl,-rpath,$ORIGIN"
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES]
    tmp = MODEL_LIB + ".tmp"
    _run(["g++"] + GXX_FLAGS + ["-fPIC", "-shared"] + inc + ["-o", tmp] + srcs +
         ["-L", HERE, "-l:libfemx.so", rpath], verbose, "libfemx_model.so")
    os.replace(tmp, MODEL_LIB)
    tmp = DRIVER + ".tmp"
    _run(["g++"] + GXX_FLAGS + inc + ["-o", tmp, os.path.join(HOST, "femx_explicit.cpp"),
                                      "-L", HERE, "-l:libfemx_model.so", "-l:libfemx.so", rpath],
         verbose, "femx_explicit")
    os.replace(tmp, DRIVER)


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    _run(cmd, verbose, "libfemx.so")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
