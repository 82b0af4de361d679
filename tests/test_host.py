"""Host-side logic: BC post-processing, generators, the C-ABI library's
exports and block partition (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from femorph_amd import femx, meshgen
from femorph_amd.mesh import BoundaryConditions, Mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "femx.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(femx_\w+)\s*\(", hdr))
    L = femx.load()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert {s for s, _, _ in femx.SIGNATURES} == declared


def test_library_refuses_without_device():
    if femx.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(femx.FemxError) as ei:
        femx.DeviceMesh(meshgen.tet_cube(2))
    assert ei.value.status == 7


def test_bc_postprocessing_mirrors_reference():
    """postProcessNodeVelocities: unique sorted nodes, tables in line order,
    one direction entry per LINE indexed per unique node (quirk)."""
    lines = [(7, 1.0, 0, 0, 0.0, 0), (3, 2.0, 0, 0, 0.0, 1), (7, 5.0, 0, 0, 1.0, 2), (3, 4.0, 0, 0, 2.0, 3)]
    bc = BoundaryConditions.from_lines(nodevel=lines, nodedirbc=[(5, 0.0), (5, 1.0, 2.0, 3.0)], nodepres=[(2, 0.5)])
    assert bc.vel_node.tolist() == [3, 7]
    assert bc.vel_row_off.tolist() == [0, 2, 4]
    assert bc.vel_rows[:, 0].tolist() == [2.0, 4.0, 1.0, 5.0]
    assert bc.vel_dir.tolist() == [0, 1]          # dirs of lines 0 and 1, not of the nodes
    assert bc.dir_val.tolist() == [[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]]
    assert bc.pres_node.tolist() == [2]


def test_generators():
    m = meshgen.tet_cube(20)
    assert (m.n_elems, m.n_nodes) == (48000, 9261)
    m = meshgen.tet_cube(5, 0.15, 3)
    x = m.coords[m.conn]
    det = np.linalg.det(np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2))
    assert np.all(det > 0) and abs(det.sum() / 6 - 1.0) < 1e-12
    h = meshgen.hex_cube(4, 0.2, 1)
    assert (h.n_elems, h.n_nodes) == (64, 125)
    c = h.coords
    onb = (np.abs(c) < 1e-12) | (np.abs(c - 1) < 1e-12)
    g = meshgen.hex_cube(4)
    assert np.array_equal(c[onb], g.coords[onb])  # no normal jitter on the boundary
    bc = meshgen.lid_cavity_bc(g)
    assert np.all(g.coords[bc.vel_node, 1] == 1.0)
    assert len(np.intersect1d(bc.vel_node, bc.dir_node)) > 0  # lid edges are walls too (walls win)
    ch = meshgen.channel_bc(g)
    assert np.all(g.coords[ch.pres_node, 0] == 1.0)


@pytest.mark.parametrize("order", [femx.BLOCKS_MORTON, femx.BLOCKS_ELEMENT_ORDER])
@pytest.mark.parametrize("kind,n,B", [("tet", 6, 256), ("tet", 9, 512), ("hex", 7, 128), ("tet", 36, 256),
                                       ("hex", 60, 128)])
def test_block_partition_invariants(kind, n, B, order):
    """The larger meshes (279,936 tets, 216,000 hexa) span several of the
    builder's independently cut ranges of 2^17 elements."""
    m = meshgen.tet_cube(n, 0.1, 1) if kind == "tet" else meshgen.hex_cube(n, 0.1, 1)
    s = femx.block_stats(m, B, order)
    assert femx.block_stats(m, B, order) == s  # deterministic (threads take ranges in any order)
    deg = np.bincount(m.conn.ravel(), minlength=m.n_nodes)
    assert s["elements"] == m.n_elems
    assert s["max_local"] <= 256 and s["max_elems"] <= B
    # every (node, block) pair is either complete (node fully inside) or one partial slot
    assert s["complete"] + s["partials"] == s["local_nodes"]
    # each node is complete in exactly one block or shared
    assert s["complete"] + s["shared_nodes"] == m.n_nodes - 0 * deg.size
    assert s["blocks"] >= (m.n_elems + B - 1) // B


def test_bc_node_out_of_range_rejected():
    m = meshgen.tet_cube(2)
    bc = BoundaryConditions(dir_node=[10 ** 6], dir_val=[[0, 0, 0]])
    assert bc.dir_node[0] == 10 ** 6 and m.n_nodes < 10 ** 6


def test_block_cap_on_corners_of_one_node():
    """Node 0 is a corner of all 512 tets: every block holds at most 128 of
    its corners, so a corner's rank in its node's run fits 7 bits."""
    m = meshgen.tet_fan(3)
    assert m.n_elems == 512 and np.bincount(m.conn.ravel())[0] == 512
    s = femx.block_stats(m, 256)
    assert s["elements"] == 512 and s["max_elems"] <= 128 and s["blocks"] >= 4
    assert s["complete"] + s["partials"] == s["local_nodes"]


def test_element_order_blocks_follow_numbering():
    """Blocks in element order keep runs of consecutive ids, so a randomly
    renumbered mesh touches many more nodes per block than in Morton order,
    while the Morton partition does not depend on the numbering."""
    m = meshgen.tet_cube(10, 0.1, 1)
    perm = np.random.default_rng(0).permutation(m.n_elems)
    r = Mesh(m.coords, m.conn[perm])
    a, b = femx.block_stats(r, 256, femx.BLOCKS_MORTON), femx.block_stats(m, 256, femx.BLOCKS_MORTON)
    assert a == b
    e = femx.block_stats(r, 256, femx.BLOCKS_ELEMENT_ORDER)
    assert e["local_nodes"] > 3 * a["local_nodes"] and e["max_local"] <= 256


def test_bad_meshes_rejected():
    """An empty mesh and connectivity outside the nodes are refused by the
    host checks, before anything is allocated (no GPU needed)."""
    from femorph_amd.mesh import Mesh
    m = meshgen.tet_cube(2)
    bad = Mesh(m.coords, m.conn.copy())
    bad.conn[3, 2] = m.n_nodes
    with pytest.raises(femx.FemxError, match="out of range"):
        femx.block_stats(bad, 256)
    neg = Mesh(m.coords, m.conn.copy())
    neg.conn[0, 0] = -1
    with pytest.raises(femx.FemxError, match="out of range"):
        femx.block_stats(neg, 256)
    empty = Mesh(m.coords, m.conn[:0].copy())
    with pytest.raises(femx.FemxError):
        femx.block_stats(empty, 256)
