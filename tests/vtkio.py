"""Minimal parser of the ASCII VTK legacy files the model class writes
(unstructured grid, POINT_DATA / CELL_DATA sections of SCALARS with a
LOOKUP_TABLE line and VECTORS).  Test helper only."""
from __future__ import annotations

import numpy as np


def read_vtk_legacy(path: str) -> dict:
    with open(path) as f:
        lines = f.read().split("\n")
    out = {"header": lines[:4], "point_data": {}, "cell_data": {}, "order": []}
    i = 4
    section = None
    n_pts = n_cells = 0
    while i < len(lines):
        ln = lines[i].strip()
        if not ln:
            i += 1
            continue
        t = ln.split()
        if t[0] == "POINTS":
            n_pts = int(t[1])
            out["points_type"] = t[2]
            out["points"] = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n_pts)])
            i += 1 + n_pts
        elif t[0] == "CELLS":
            n_cells, size = int(t[1]), int(t[2])
            cells = [[int(v) for v in lines[i + 1 + k].split()] for k in range(n_cells)]
            assert sum(len(c) for c in cells) == size
            out["cells"] = [c[1:] for c in cells]
            assert all(c[0] == len(c) - 1 for c in cells)
            i += 1 + n_cells
        elif t[0] == "CELL_TYPES":
            out["cell_types"] = np.array([int(lines[i + 1 + k]) for k in range(int(t[1]))])
            i += 1 + int(t[1])
        elif t[0] in ("POINT_DATA", "CELL_DATA"):
            section = "point_data" if t[0] == "POINT_DATA" else "cell_data"
            assert int(t[1]) == (n_pts if section == "point_data" else n_cells)
            i += 1
        elif t[0] in ("SCALARS", "VECTORS"):
            name, n = t[1], (n_pts if section == "point_data" else n_cells)
            j = i + 1
            if t[0] == "SCALARS":
                assert lines[j].strip() == "LOOKUP_TABLE default"
                j += 1
            vals = np.array([[float(v) for v in lines[j + k].split()] for k in range(n)])
            out[section][name] = vals
            out["order"].append((section, t[0], name, " ".join(t[2:])))
            i = j + n
        else:
            raise ValueError(f"unexpected VTK line {i + 1}: {ln!r}")
    return out
