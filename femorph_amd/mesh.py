"""Host-side mesh and boundary-condition containers.

They mirror the public data members of the reference's ``femModel`` that the
explicit solvers consume (src/femModel.h:22-108):

* ``nodeList[i]->coords``                 -> ``Mesh.coords``  [N, 3] float64
* ``elementList[e]->elementConnections``  -> ``Mesh.conn``    [E, npe] int32
* ``velNodesID / velNodesDir / velNodesTimeVals`` (after
  ``postProcessNodeVelocities``, src/femModel.cpp:4549-4576)
                                          -> ``BoundaryConditions.vel_*``
* ``diricheletBCNode / diricheletBCValues`` -> ``dir_node / dir_val``
* ``presNodesID / presNodesValue``          -> ``pres_node / pres_val``
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    coords: np.ndarray  # [N, 3] float64
    conn: np.ndarray  # [E, npe] int32, 0-based

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(self.conn, dtype=np.int32)
        if self.conn.ndim != 2 or self.conn.shape[1] not in (4, 8):
            raise ValueError("connectivity must be [E, 4] (tet4) or [E, 8] (hexa8)")
        if self.conn.size and (self.conn.min() < 0 or self.conn.max() >= len(self.coords)):
            raise ValueError("connectivity references a node outside the mesh")

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.conn.shape[0])

    @property
    def npe(self) -> int:
        return int(self.conn.shape[1])


@dataclass
class BoundaryConditions:
    """Boundary conditions in the reference's post-processed form.

    ``vel_node`` holds the unique, sorted prescribed-velocity nodes;
    ``vel_rows[vel_row_off[i]:vel_row_off[i+1]]`` is node i's time table of
    rows (vx, vy, vz, t) in file order; ``vel_dir[i]`` is the direction code
    of the i-th NODEVEL *line* (0..6 = X, Y, Z, XY, YZ, XZ, XYZ), which is how
    the reference indexes it (src/femModel.cpp:4570 vs :4590).
    """

    vel_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    vel_dir: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    vel_row_off: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    vel_rows: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    dir_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dir_val: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    pres_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pres_val: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __post_init__(self):
        self.vel_node = np.ascontiguousarray(self.vel_node, np.int32)
        self.vel_dir = np.ascontiguousarray(self.vel_dir, np.int32)
        self.vel_row_off = np.ascontiguousarray(self.vel_row_off, np.int32)
        self.vel_rows = np.ascontiguousarray(self.vel_rows, np.float64).reshape(-1, 4)
        self.dir_node = np.ascontiguousarray(self.dir_node, np.int32)
        self.dir_val = np.ascontiguousarray(self.dir_val, np.float64).reshape(-1, 3)
        self.pres_node = np.ascontiguousarray(self.pres_node, np.int32)
        self.pres_val = np.ascontiguousarray(self.pres_val, np.float64)

    @staticmethod
    def from_lines(nodevel=(), nodedirbc=(), nodepres=()):
        """Build from NODEVEL / NODEDIRBC / NODEPRES records (0-based nodes).

        ``nodevel``: iterable of (node, vx, vy, vz, t[, dir]) in file order;
        post-processed like ``femModel::postProcessNodeVelocities``: unique
        sorted node ids (``femUtils::MakeCompactList``), per-node tables in
        line order, one direction entry per line.
        """
        nv = [tuple(r) for r in nodevel]
        if nv:
            nodes = np.array([int(r[0]) for r in nv], np.int64)
            uniq = np.unique(nodes)
            idx = np.searchsorted(uniq, nodes)
            order = np.argsort(idx, kind="stable")
            rows = np.array([[r[1], r[2], r[3], r[4]] for r in nv], np.float64)[order]
            counts = np.bincount(idx, minlength=len(uniq))
            off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
            dirs = np.array([int(r[5]) if len(r) > 5 else 6 for r in nv], np.int32)[: len(uniq)]
            vel = dict(vel_node=uniq.astype(np.int32), vel_dir=dirs, vel_row_off=off, vel_rows=rows)
        else:
            vel = {}
        dn = [tuple(r) for r in nodedirbc]
        dirb = {}
        if dn:
            dirb = dict(
                dir_node=np.array([int(r[0]) for r in dn], np.int32),
                dir_val=np.array([[r[1], r[2] if len(r) > 2 else 0.0, r[3] if len(r) > 3 else 0.0] for r in dn]),
            )
        pn = [tuple(r) for r in nodepres]
        pres = {}
        if pn:
            pres = dict(pres_node=np.array([int(r[0]) for r in pn], np.int32),
                        pres_val=np.array([float(r[1]) for r in pn]))
        return BoundaryConditions(**vel, **dirb, **pres)

    @property
    def n_vel(self) -> int:
        return int(len(self.vel_node))
