#!/usr/bin/env python3
"""Writes tests/golden/ldc_coarse_model.txt.gz from the reference's example
02_lid_driven_cavity_coarse (a data file: 1,118 NODE, 4,444 TET4 ELEMENT,
298 NODEVEL, 635 NODEDIRBC lines plus NODEDOF / PROBLEM / TIMESTEP /
VMSPROPS), byte for byte, gzip-compressed.  It is the reader's real-input
fixture: the C++ and Python readers must agree on it and the VMS step must
run on it.

Run in the build container (the reference is not present on the GPU box):
  python tests/golden/make_model_fixture.py
"""
import gzip
import os
import sys

REF = os.environ.get("FEMX_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "examples/02_explicit_fluid_VMS/02_lid_driven_cavity_coarse/model.txt")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ldc_coarse_model.txt.gz")


def main():
    data = open(SRC, "rb").read()
    with gzip.GzipFile(OUT, "wb", mtime=0) as f:
        f.write(data)
    print(f"wrote {OUT}: {len(data)} bytes of model.txt")


if __name__ == "__main__":
    sys.exit(main())
