#!/usr/bin/env python3
"""Writes the reference's example VMS inputs as gzip-compressed fixtures under
tests/golden/, byte for byte (data files, not source):

- ldc_coarse_model.txt.gz from 02_lid_driven_cavity_coarse (1,118 NODE,
  4,444 TET4 ELEMENT, 298 NODEVEL, 635 NODEDIRBC lines plus NODEDOF /
  PROBLEM / TIMESTEP / VMSPROPS);
- cylinder_model.txt.gz from 01_cylinder (8,595 NODE, 45,754 TET4 ELEMENT,
  202 NODEVEL rows ramping the inflow from t = 0 to t = 1, 2,565 NODEDIRBC,
  102 NODEPRES outlet nodes, DUALSTEP 1e-3 0);
- ldc_medium_model.txt.gz from 03_lid_driven_cavity_medium (7,309 NODE,
  36,468 TET4 ELEMENT, 1,022 NODEVEL, 2,392 NODEDIRBC, DUALSTEP 1e-6 100).

The readers must agree on them and the VMS step runs on them against the
oracle.

Run in the build container (the reference is not present on the GPU box):
  python tests/golden/make_model_fixture.py
"""
import gzip
import os
import sys

REF = os.environ.get("FEMX_REFERENCE", "/root/reference")
EX = os.path.join(REF, "examples/02_explicit_fluid_VMS")
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = {
    "ldc_coarse_model.txt.gz": "02_lid_driven_cavity_coarse/model.txt",
    "cylinder_model.txt.gz": "01_cylinder/model.txt",
    "ldc_medium_model.txt.gz": "03_lid_driven_cavity_medium/model.txt",
}


def main():
    for out, src in FIXTURES.items():
        data = open(os.path.join(EX, src), "rb").read()
        path = os.path.join(HERE, out)
        with gzip.GzipFile(path, "wb", compresslevel=9, mtime=0) as f:
            f.write(data)
        print(f"wrote {path}: {len(data)} bytes of {src}")


if __name__ == "__main__":
    sys.exit(main())
