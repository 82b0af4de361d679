#!/usr/bin/env python3
"""Writes tests/golden/ldc_mesh2.npz from the reference's example
08_2D_hexa_ldc_mesh2 (data files only):

- fluid.vtu: the 29 x 29 x 1 hexa8 lid-driven cavity mesh (1,800 nodes,
  841 elements), node coordinates and the connectivity, in file order;
- wall_top.vtp, wall_left.vtp, wall_right.vtp, wall_bottom.vtp: the
  GlobalNodeID (1-based) of the lid and wall nodes, which its model writer
  (create_exVMS_simul.py) turns into NODEVEL (lid u = 1) and NODEDIRBC
  (walls, zero) lines;
- 07_comparison/{hor05,vert05}_{100,1K,10K}.csv: velocity sampled by the
  reference's own Re = 100, 1,000 and 10,000 runs along the vertical
  (x = 0.5) and horizontal (y = 0.5) centrelines, columns DXYZ:0..2,
  velocity:0..2, dirBC, neuBC, Points:0..2 (the fields of its VTK writer);
- 07_comparison/re_100_horizontal.txt and re_100_vertical.txt: the
  published Re = 100 centreline benchmark values the reference plots its run
  against.

The .vtu/.vtp files are VTK XML with zlib-compressed, base64-encoded
appended arrays (UInt32 block headers); they are decoded here with zlib and
base64 only.

Run in the build container (the reference is not present on the GPU box):
  python tests/golden/make_ldc_mesh2_fixture.py
"""
import base64
import os
import re
import struct
import sys
import zlib

import numpy as np

REF = os.environ.get("FEMX_REFERENCE", "/root/reference")
EX = os.path.join(REF, "examples/02_explicit_fluid_VMS/08_2D_hexa_ldc_mesh2")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ldc_mesh2.npz")
DTYPES = {"Int64": "<i8", "Int32": "<i4", "UInt8": "u1", "Float64": "<f8", "Float32": "<f4"}


def _b64len(nbytes):
    return (nbytes + 2) // 3 * 4


def read_vtk_xml(path):
    """{name: array} of every appended DataArray of a VTK XML file
    (header_type UInt32, vtkZLibDataCompressor, base64 encoding)."""
    text = open(path, "rb").read().decode("latin-1")
    assert 'header_type="UInt32"' in text and "vtkZLibDataCompressor" in text
    m = re.search(r'<AppendedData encoding="base64">\s*_', text)
    blob = text[m.end():]
    out = {}
    for tag in re.findall(r"<DataArray[^>]*>", text):
        name = re.search(r'Name="([^"]+)"', tag).group(1)
        dt = np.dtype(DTYPES[re.search(r'type="([^"]+)"', tag).group(1)])
        ncomp = int(re.search(r'NumberOfComponents="(\d+)"', tag).group(1)) if "NumberOfComponents" in tag else 1
        off = int(re.search(r'offset="(\d+)"', tag).group(1))
        nblocks = struct.unpack("<I", base64.b64decode(blob[off:off + 8])[:4])[0]
        hlen = _b64len(4 * (3 + nblocks))
        hdr = struct.unpack(f"<{3 + nblocks}I", base64.b64decode(blob[off:off + hlen])[:4 * (3 + nblocks)])
        csizes = hdr[3:]
        data = base64.b64decode(blob[off + hlen:off + hlen + _b64len(sum(csizes))])
        raw, pos = b"", 0
        for c in csizes:
            raw += zlib.decompress(data[pos:pos + c])
            pos += c
        arr = np.frombuffer(raw, dt)
        if ncomp > 1:
            arr = arr.reshape(-1, ncomp)
        # several arrays can share a name (the cell blocks of a .vtp): keep the first
        out.setdefault(name, arr)
    return out


def main():
    vtu = read_vtk_xml(os.path.join(EX, "fluid.vtu"))
    coords = vtu["Points"].astype(np.float64)
    conn = vtu["connectivity"].reshape(-1, 8).astype(np.int32)  # 0-based, 8 per hexa8
    assert (np.diff(vtu["offsets"]) == 8).all() and (vtu["types"] == 12).all()  # VTK_HEXAHEDRON
    gid = vtu["GlobalNodeID"]
    assert (gid == np.arange(1, len(coords) + 1)).all()  # node i is NODE i + 1
    walls = {}
    for w in ("top", "left", "right", "bottom"):
        walls[w] = read_vtk_xml(os.path.join(EX, f"wall_{w}.vtp"))["GlobalNodeID"].astype(np.int32)
    cmp = os.path.join(EX, "07_comparison")
    hor = np.loadtxt(os.path.join(cmp, "hor05_100.csv"), skiprows=1, delimiter=",")
    ver = np.loadtxt(os.path.join(cmp, "vert05_100.csv"), skiprows=1, delimiter=",")
    bench_h = np.loadtxt(os.path.join(cmp, "re_100_horizontal.txt"), skiprows=1)
    bench_v = np.loadtxt(os.path.join(cmp, "re_100_vertical.txt"), skiprows=1)
    more = {}
    for re in ("1K", "10K"):
        for line in ("hor05", "vert05"):
            more[f"{line}_{re}"] = np.loadtxt(os.path.join(cmp, f"{line}_{re}.csv"), skiprows=1, delimiter=",")
    np.savez_compressed(OUT, coords=coords, conn=conn, top=walls["top"], left=walls["left"],
                        right=walls["right"], bottom=walls["bottom"], hor05_100=hor, vert05_100=ver,
                        bench_horizontal=bench_h, bench_vertical=bench_v, **more)
    print(f"wrote {OUT}: {len(coords)} nodes, {len(conn)} hexa8, lid {len(walls['top'])} nodes, "
          f"probes {len(hor)} + {len(ver)}")


if __name__ == "__main__":
    sys.exit(main())
