"""feMorph's ``model.txt`` format: reader and writer.

The reader follows the grammar of ``femModel::ReadFromFEMTextFile``
(src/femModel.cpp:3719-4213): each line is trimmed and split on any of
``' '`` and ``','`` with runs of separators compressed; the first token is a
case-insensitive keyword; indices in the file are 1-based; the node index is
the order of the ``NODE`` lines.  Prescribed velocities are post-processed
like ``postProcessNodeVelocities`` (src/femModel.cpp:4549-4576).  The C++
drop-in model class (femorph_amd/csrc/host/femModel.cpp) reads the same
files; tests check that both readers agree.

The writer produces files both readers accept, so generated workloads, the
GPU path and the oracle consume identical inputs.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import BoundaryConditions, Mesh

NODES_OF_TYPE = {"ROD": 2, "TRI3": 3, "QUAD4": 4, "TET4": 4, "TET10": 10, "HEXA8": 8}
FIXED_DOFS = ("NONE", "X", "Y", "Z", "XY", "YZ", "XZ")
PROBLEMS = {"POISSON": 0, "PPE": 1, "DISTANCE": 2, "EXPLICITVMS": 3}


class ModelFormatError(ValueError):
    pass


@dataclass
class Model:
    """The fields of femModel the explicit and implicit solvers consume
    (src/femModel.h:22-108)."""

    coords: np.ndarray
    elem_types: list
    conn: list  # per element, 0-based node ids
    bc: BoundaryConditions
    vmsprops: list = field(default_factory=list)
    exopts: list = field(default_factory=list)
    max_node_dofs: int = 0
    fixed_dofs: str = ""
    time_step: float = 0.0
    total_steps: int = 0
    save_every: int = 0
    dual_step: float = 0.0
    tot_dual_steps: int = 0
    exit_condition: float = 0.0
    problem: int = 0
    output: str = ""
    log_file: str = ""
    # implicit solver (femIncompressibleSolver) fields, src/femModel.h
    alpha_m: float = 0.0
    alpha_f: float = 0.0
    gamma: float = 0.0
    stages: list = field(default_factory=list)
    prescribed_vel: int = -1  # PRESCRIBEDVEL type, -1 when absent
    ini: list = field(default_factory=list)  # (node 0-based, dof, value)
    el_velocity: list = field(default_factory=list)  # ELVELS rows in file order
    el_diffusivity: list = field(default_factory=list)  # ELDIFF rows in file order

    def mesh(self) -> Mesh:
        """The element list as a single-type Mesh (tet4 or hexa8)."""
        types = set(self.elem_types)
        if len(types) != 1 or next(iter(types)) not in ("TET4", "HEXA8"):
            raise ModelFormatError(f"explicit solvers need one element type, TET4 or HEXA8; got {sorted(types)}")
        return Mesh(self.coords, np.array(self.conn, np.int32))


def _tokens(line: str):
    out, cur, sep = [], [], False
    for ch in line.strip():
        if ch in " ,":
            if not sep:
                out.append("".join(cur))
                cur, sep = [], True
        else:
            cur.append(ch)
            sep = False
    out.append("".join(cur))
    return out


def _f(t, i, what):
    try:
        return float(t[i])
    except (IndexError, ValueError):
        raise ModelFormatError(f"Invalid {what} Format") from None


def _i(t, i, what):
    try:
        return int(t[i])
    except (IndexError, ValueError):
        raise ModelFormatError(f"Invalid {what} Format") from None


def read_model_txt(path: str) -> Model:
    coords, types, conn = [], [], []
    nodevel, nodedir, nodepres = [], [], []
    m = dict(vmsprops=[], exopts=[])
    with open(path) as fh:
        for line in fh:
            t = _tokens(line)
            key = t[0].upper()
            if key == "NODE":
                coords.append([_f(t, 2, "Node"), _f(t, 3, "Node"), _f(t, 4, "Node")])
            elif key == "ELEMENT":
                ty = t[1].strip().upper() if len(t) > 1 else ""
                nn = NODES_OF_TYPE.get(ty)
                if nn is None:
                    raise ModelFormatError("Invalid ELEMENT Format")
                types.append(ty)
                conn.append([_i(t, 4 + a, "ELEMENT") - 1 for a in range(nn)])
            elif key == "NODEDIRBC":
                n = _i(t, 1, "NODEDIRBC") - 1
                vx = _f(t, 2, "NODEDIRBC")
                vy = vz = 0.0
                if len(t) > 3:
                    vy, vz = _f(t, 3, "NODEDIRBC"), _f(t, 4, "NODEDIRBC")
                nodedir.append((n, vx, vy, vz))
            elif key == "NODEVEL":
                row = [_i(t, 1, "NODEVEL") - 1] + [_f(t, k, "NODEVEL") for k in (2, 3, 4, 5)]
                row.append(int(_f(t, 6, "NODEVEL")) if len(t) > 6 else 6)
                nodevel.append(tuple(row))
            elif key == "NODEPRES":
                nodepres.append((_i(t, 1, "NODEPRES") - 1, _f(t, 2, "NODEPRES")))
            elif key == "NODEDOF":
                m["max_node_dofs"] = _i(t, 1, "NODEDOF")
                if len(t) > 2:
                    fd = t[2].strip().upper()
                    if fd not in FIXED_DOFS:
                        raise ModelFormatError("Invalid NODEDOF Option Format")
                    m["fixed_dofs"] = fd
            elif key == "PROBLEM":
                p = t[1].strip().upper() if len(t) > 1 else ""
                if p not in PROBLEMS:
                    raise ModelFormatError("Invalid PROBLEM Type")
                m["problem"] = PROBLEMS[p]
            elif key == "OUTPUT":
                m["output"] = t[1].strip() if len(t) > 1 else ""
                if not m["output"]:
                    raise ModelFormatError("Invalid solution file prefix")
            elif key == "LOGFILE":
                m["log_file"] = t[1].strip() if len(t) > 1 else ""
                if not m["log_file"]:
                    raise ModelFormatError("Invalid log file name")
            elif key == "TIMESTEP":
                m["time_step"] = _f(t, 1, "TIMESTEP")
                m["total_steps"] = _i(t, 2, "TIMESTEP")
                m["save_every"] = _i(t, 3, "TIMESTEP")
            elif key == "DUALSTEP":
                m["dual_step"] = _f(t, 1, "TIMESTEP")
                m["tot_dual_steps"] = _i(t, 2, "TIMESTEP")
            elif key == "EXITCOND":
                m["exit_condition"] = _f(t, 1, "EXITCOND")
            elif key == "VMSPROPS":
                m["vmsprops"] += [_f(t, k, "VMSPROPS") for k in range(1, 6)]
            elif key == "EXOPTS":
                m["exopts"] += [1.0 if _i(t, k, "EXOPTS") != 0 else 0.0 for k in range(1, 7)]
            elif key == "TIMEINTEGRATION":
                m["alpha_m"], m["alpha_f"], m["gamma"] = (_f(t, k, "TIMEINTEGRATION") for k in (1, 2, 3))
            elif key == "STAGES":
                m["stages"] = [_i(t, k, "STAGES") for k in range(1, len(t))]
            elif key == "PRESCRIBEDVEL":
                m["prescribed_vel"] = int(_f(t, 1, "PRESCRIBEDVEL"))
            elif key == "INI":
                m.setdefault("ini", []).append((_i(t, 1, "INI") - 1, _i(t, 2, "INI"), _f(t, 3, "INI")))
            elif key == "ELVELS":
                m.setdefault("el_velocity", []).append([_f(t, k, "Element Velocity") for k in (2, 3, 4)])
            elif key == "ELDIFF":
                m.setdefault("el_diffusivity", []).append([_f(t, k, "Element Diffusivity") for k in (2, 3, 4)])
    bc = BoundaryConditions.from_lines(nodevel, nodedir, nodepres)
    # one direction entry per NODEVEL line, indexed by unique node afterwards
    if nodevel:
        bc.vel_dir = np.array([r[5] for r in nodevel], np.int32)[: len(bc.vel_node)]
    return Model(np.array(coords, np.float64).reshape(-1, 3), types, conn, bc, **m)


def write_model_txt(path: str, mesh: Mesh, bc: BoundaryConditions, *, problem="EXPLICITVMS",
                    time_step=0.01, total_steps=2, save_every=1, dual_step=None, tot_dual_steps=0,
                    vmsprops=(1.0, 0.01, 1.0, 4.0, 2.0), exopts=None, max_node_dofs=4, fixed_dofs="NONE",
                    exit_condition=None, output=None, log_file=None, time_integration=None, stages=None,
                    prescribed_vel=None, ini=(), el_velocity=None, el_diffusivity=None) -> None:
    """Write a model file.  Values are written with 17 significant digits, so
    both readers recover the arrays exactly.  NODEVEL lines are grouped per
    node; the reference indexes their direction codes by line, so a model
    whose nodes carry several rows must use a single direction code."""
    nv = bc.vel_row_off[1:] - bc.vel_row_off[:-1] if len(bc.vel_node) else np.zeros(0, np.int64)
    if len(bc.vel_node) and nv.max() > 1 and len(set(bc.vel_dir.tolist())) > 1:
        raise ValueError("several velocity rows per node need a single direction code")
    etype = {4: "TET4", 8: "HEXA8"}[mesh.npe]
    g = "%.17g"
    with open(path, "w") as f:
        for i, x in enumerate(mesh.coords):
            f.write(f"NODE {i + 1} {g % x[0]} {g % x[1]} {g % x[2]}\n")
        for e, c in enumerate(mesh.conn):
            f.write(f"ELEMENT {etype} {e + 1} 1 " + " ".join(str(int(k) + 1) for k in c) + "\n")
        for i, n in enumerate(bc.vel_node):
            d = int(bc.vel_dir[i]) if i < len(bc.vel_dir) else 6
            for r in bc.vel_rows[bc.vel_row_off[i]:bc.vel_row_off[i + 1]]:
                f.write(f"NODEVEL {int(n) + 1} {g % r[0]} {g % r[1]} {g % r[2]} {g % r[3]} {d}\n")
        for n, v in zip(bc.dir_node, bc.dir_val):
            f.write(f"NODEDIRBC {int(n) + 1} {g % v[0]} {g % v[1]} {g % v[2]}\n")
        for n, v in zip(bc.pres_node, bc.pres_val):
            f.write(f"NODEPRES {int(n) + 1} {g % v}\n")
        f.write(f"NODEDOF {max_node_dofs} {fixed_dofs}\n")
        f.write(f"PROBLEM {problem}\n")
        f.write(f"TIMESTEP {g % time_step} {total_steps} {save_every}\n")
        f.write(f"DUALSTEP {g % (time_step if dual_step is None else dual_step)} {tot_dual_steps}\n")
        # both solvers' readers take five values (SUPG uses the first four)
        vp = list(vmsprops) + [0.0] * (5 - len(vmsprops))
        f.write("VMSPROPS " + " ".join(g % v for v in vp) + "\n")
        if exopts is not None:
            f.write("EXOPTS " + " ".join(str(int(v)) for v in exopts) + "\n")
        if exit_condition is not None:
            f.write(f"EXITCOND {g % exit_condition}\n")
        if output:
            f.write(f"OUTPUT {output}\n")
        if log_file:
            f.write(f"LOGFILE {log_file}\n")
        if time_integration is not None:
            f.write("TIMEINTEGRATION " + " ".join(g % v for v in time_integration) + "\n")
        if stages is not None:
            f.write("STAGES " + " ".join(str(int(v)) for v in stages) + "\n")
        if prescribed_vel is not None:
            f.write(f"PRESCRIBEDVEL {int(prescribed_vel)}\n")
        for n, dof, v in ini:
            f.write(f"INI {int(n) + 1} {int(dof)} {g % v}\n")
        for key, rows in (("ELVELS", el_velocity), ("ELDIFF", el_diffusivity)):
            if rows is not None:
                for e, r in enumerate(np.asarray(rows, np.float64).reshape(-1, 3)):
                    f.write(f"{key} {e + 1} {g % r[0]} {g % r[1]} {g % r[2]}\n")
