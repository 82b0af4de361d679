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
OBJ_DIR = os.path.join(HERE, "_build")

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
DROPIN = os.path.join(CSRC, "dropin")
MODEL_LIB = os.path.join(HERE, "libfemx_model.so")
DRIVER = os.path.join(HERE, "femx_explicit")
# the standalone feMorph model (femModel, femSolver base, femPointGrid) ...
HOST_SOURCES = ["femModel.cpp", "femSolver.cpp", "femPointGrid.cpp"]
# ... and the drop-in solver classes, which compile against it here and
# against the reference's own headers inside a feMorph build
DROPIN_SOURCES = ["femVMSExplicitFluidSolver.cpp", "femSUPGExplicitFluidSolver.cpp", "femIncompressibleSolver.cpp"]
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
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES] + [os.path.join(DROPIN, s) for s in DROPIN_SOURCES]
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
    """Each .hip source compiles to an object in parallel (hipcc, gfx950),
    then one link makes libfemx.so; the host libraries follow."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    os.makedirs(OBJ_DIR, exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    inc = ["-I", os.path.join(ROOT, "include")]
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in srcs]
    jobs = int(os.environ.get("MAX_JOBS", "0") or 0) or min(len(srcs), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, 16))) as ex:
        list(ex.map(lambda so: _run([hipcc()] + flags + inc + ["-c", so[0], "-o", so[1]], verbose, so[0]),
                    zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([hipcc()] + HIPCC_FLAGS + ["-o", tmp] + objs, verbose, "libfemx.so")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
