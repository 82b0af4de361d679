#!/usr/bin/env python3
"""Writes tests/golden/lumped_mass_06.npz from the reference's example 06.

Inputs (data files of the reference, read as data only):
  examples/02_explicit_fluid_VMS/06_2D_hexa_lid_driven_cavity/fluid.vtu
      the 100 x 100 x 1 hexa8 mesh (20,000 nodes, 9,801 HEXA8) that the
      example's create_exVMS_simul.py turns into model.txt, writing node
      coordinates with '%.8e' and connectivity in VTK cell order;
  examples/02_explicit_fluid_VMS/06_2D_hexa_lid_driven_cavity/lumped_mass.txt
      form_lumped_mass of that model with rho = 10 (VMSPROPS 10.0 ...),
      written by femUtils::writeVectorToFile ('%25.10e', one value per line,
      l_M[node*3 + dof]).
Output: coords (rounded like model.txt), conn (0-based), rho, expected l_M.

Run in the build container (the reference is not present on the GPU box):
  python tests/golden/make_lumped_mass_fixture.py
"""
import base64
import os
import re
import struct
import sys
import zlib

import numpy as np

REF = os.environ.get("FEMX_REFERENCE", "/root/reference")
EX = os.path.join(REF, "examples/02_explicit_fluid_VMS/06_2D_hexa_lid_driven_cavity")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lumped_mass_06.npz")


def read_vtu_appended_base64(path):
    raw = open(path, "rb").read()
    head = raw[: raw.index(b"<AppendedData")].decode()
    arrays = re.findall(r'<DataArray type="(\w+)"[^>]*?Name="(\w+)"[^>]*?offset="(\d+)"', head)
    start = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    dt = {"Int64": "<i8", "Int32": "<i4", "Float64": "<f8", "UInt8": "u1"}
    out = {}
    for typ, name, off in arrays:
        p = start + int(off)
        first = base64.b64decode(raw[p: p + 16])
        nb = struct.unpack("<I", first[:4])[0]
        hlen = 4 * (3 + nb)
        henc = ((hlen + 2) // 3) * 4
        hdr = struct.unpack("<%dI" % (3 + nb), base64.b64decode(raw[p: p + henc])[:hlen])
        comp = hdr[3:]
        denc = ((sum(comp) + 2) // 3) * 4
        data = base64.b64decode(raw[p + henc: p + henc + denc])
        buf, q = b"", 0
        for c in comp:
            buf += zlib.decompress(data[q: q + c])
            q += c
        out[name] = np.frombuffer(buf, dtype=dt[typ])
    return out


def main():
    d = read_vtu_appended_base64(os.path.join(EX, "fluid.vtu"))
    pts = d["Points"].reshape(-1, 3)
    # create_exVMS_simul.py writes 'NODE %d %.8e %.8e %.8e'
    coords = np.array([[float("%.8e" % v) for v in row] for row in pts])
    off = d["offsets"]
    assert np.all(np.diff(np.concatenate([[0], off])) == 8) and np.all(d["types"] == 12)
    conn = d["connectivity"].reshape(-1, 8).astype(np.int32)
    lm = np.loadtxt(os.path.join(EX, "lumped_mass.txt"))
    assert lm.shape == (3 * len(coords),)
    np.savez_compressed(OUT, coords=coords, conn=conn, rho=np.float64(10.0), l_M=lm)
    print(f"wrote {OUT}: {len(coords)} nodes, {len(conn)} hexa8, {len(lm)} values")


if __name__ == "__main__":
    sys.exit(main())
