"""Seeded synthetic meshes and boundary conditions for the explicit solvers.

The reference ships only a handful of example meshes; the workloads here are
generated on the host, identically for the GPU path and the CPU oracle:

* ``tet_cube``  : n^3 structured hexes, each split into 6 positively oriented
  tet4 (Kuhn split along the main diagonal), optional node jitter;
* ``hex_cube``  : n^3 hexa8 in the reference's local node order
  (src/femHexa8.cpp:15-17), optional node jitter;
* ``lid_cavity_bc``: lid y = 1 moving at (U, 0, 0) through a NODEVEL time
  table [0, 1e9], every other face NODEDIRBC 0 (walls win on shared edge
  nodes, as in the reference's BC order velocity -> Dirichlet);
* ``channel_bc``: parabolic inflow at x = 0, outflow p = 0 at x = 1, no-slip
  walls on y and z faces.

Jitter moves each node component by U(-j h, j h), tangentially only on the
boundary (a component normal to a boundary face is never moved), so the
domain stays the unit cube.
"""
from __future__ import annotations

import itertools

import numpy as np

from .mesh import BoundaryConditions, Mesh

# Kuhn split: one tet per permutation of the axes, walking v000 -> v111.
_KUHN = []
for perm in itertools.permutations(range(3)):
    v = [0, 0, 0]
    verts = [tuple(v)]
    for ax in perm:
        v[ax] = 1
        verts.append(tuple(v))
    # orientation of (v1-v0, v2-v0, v3-v0) equals the permutation parity
    m = np.array([np.subtract(verts[k], verts[0]) for k in (1, 2, 3)], float)
    if np.linalg.det(m) < 0:
        verts[1], verts[2] = verts[2], verts[1]
    _KUHN.append(verts)
_HEX_LOCAL = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _grid_nodes(n: int, jitter: float, seed: int) -> np.ndarray:
    g = np.linspace(0.0, 1.0, n + 1)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    coords = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        h = 1.0 / n
        d = rng.uniform(-jitter * h, jitter * h, size=coords.shape)
        on_bnd = (np.abs(coords) < 1e-12) | (np.abs(coords - 1.0) < 1e-12)
        d[on_bnd] = 0.0
        coords = coords + d
    return coords


def _node_id(n, i, j, k):
    return i + (n + 1) * (j + (n + 1) * k)


def tet_cube(n: int, jitter: float = 0.0, seed: int = 0) -> Mesh:
    coords = _grid_nodes(n, jitter, seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    tets = []
    for verts in _KUHN:
        tets.append(np.stack([_node_id(n, i + a, j + b, k + c) for (a, b, c) in verts], axis=1))
    conn = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    mesh = Mesh(coords, conn)
    det = _tet_det(mesh)
    if not np.all(det > 0):
        raise ValueError("jitter produced inverted tets")
    return mesh


def hex_cube(n: int, jitter: float = 0.0, seed: int = 0) -> Mesh:
    coords = _grid_nodes(n, jitter, seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    conn = np.stack([_node_id(n, i + a, j + b, k + c) for (a, b, c) in _HEX_LOCAL], axis=1).astype(np.int32)
    return Mesh(coords, conn)


def _faces(coords, tol=1e-12):
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    return dict(x0=np.abs(x) < tol, x1=np.abs(x - 1) < tol, y0=np.abs(y) < tol, y1=np.abs(y - 1) < tol,
                z0=np.abs(z) < tol, z1=np.abs(z - 1) < tol)


def lid_cavity_bc(mesh: Mesh, u_lid: float = 1.0, t_end: float = 1.0e9) -> BoundaryConditions:
    f = _faces(mesh.coords)
    lid = np.nonzero(f["y1"])[0].astype(np.int32)
    walls = np.nonzero(f["x0"] | f["x1"] | f["y0"] | f["z0"] | f["z1"])[0].astype(np.int32)
    rows = np.zeros((2 * len(lid), 4))
    rows[0::2] = [u_lid, 0.0, 0.0, 0.0]
    rows[1::2] = [u_lid, 0.0, 0.0, t_end]
    return BoundaryConditions(
        vel_node=lid, vel_dir=np.full(len(lid), 6, np.int32),
        vel_row_off=np.arange(0, 2 * len(lid) + 1, 2, dtype=np.int32), vel_rows=rows,
        dir_node=walls, dir_val=np.zeros((len(walls), 3)))


def channel_bc(mesh: Mesh, u_max: float = 1.0, t_end: float = 1.0e9) -> BoundaryConditions:
    f = _faces(mesh.coords)
    inflow = np.nonzero(f["x0"])[0].astype(np.int32)
    outflow = np.nonzero(f["x1"])[0].astype(np.int32)
    walls = np.nonzero(f["y0"] | f["y1"] | f["z0"] | f["z1"])[0].astype(np.int32)
    y, z = mesh.coords[inflow, 1], mesh.coords[inflow, 2]
    u = 16.0 * u_max * y * (1.0 - y) * z * (1.0 - z)
    rows = np.zeros((2 * len(inflow), 4))
    rows[0::2, 0] = u
    rows[1::2, 0] = u
    rows[1::2, 3] = t_end
    return BoundaryConditions(
        vel_node=inflow, vel_dir=np.full(len(inflow), 6, np.int32),
        vel_row_off=np.arange(0, 2 * len(inflow) + 1, 2, dtype=np.int32), vel_rows=rows,
        dir_node=walls, dir_val=np.zeros((len(walls), 3)),
        pres_node=outflow, pres_val=np.zeros(len(outflow)))


def random_velocity(mesh: Mesh, amp: float = 1.0e-3, seed: int = 0, ndof: int = 4) -> np.ndarray:
    """Initial state: each velocity component U(-amp, amp) per node, p = 0."""
    rng = np.random.default_rng(seed)
    sol = np.zeros((mesh.n_nodes, ndof))
    sol[:, :3] = rng.uniform(-amp, amp, size=(mesh.n_nodes, 3))
    return sol


def _tet_det(mesh: Mesh) -> np.ndarray:
    """6 V of every tet: (x1 - x0) . ((x2 - x0) x (x3 - x0))."""
    x = mesh.coords[mesh.conn]
    a, b, c = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]
    return np.einsum("ij,ij->i", a, np.cross(b, c))


def min_tet_h(mesh: Mesh) -> float:
    """min over elements of (6 V)^(1/3), the reference's VMS length scale."""
    return float(np.min(np.cbrt(_tet_det(mesh))))


def tet_fan(level: int = 3) -> Mesh:
    """Tets (centre, face) over an octahedron subdivided `level` times and
    projected to the unit sphere: node 0 is a corner of every one of the
    8 * 4**level tets, far above what one element block may hold of one node."""
    verts = [np.array(v, float) for v in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))]
    faces = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    for _ in range(level):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                v = verts[i] + verts[j]
                verts.append(v / np.linalg.norm(v))
                cache[key] = len(verts) - 1
            return cache[key]

        new = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new
    coords = np.vstack([np.zeros(3)] + verts)
    conn = np.array([(0, a + 1, b + 1, c + 1) for a, b, c in faces], np.int32)
    m = Mesh(coords, conn)
    neg = _tet_det(m) < 0
    conn[neg, 2], conn[neg, 3] = conn[neg, 3].copy(), conn[neg, 2].copy()
    return Mesh(coords, conn)
