"""model.txt reader/writer and VTK legacy writer of the drop-in model class
(femorph_amd/csrc/host/femModel.cpp), on the CPU: the driver's model-only
modes read and export without touching a GPU.

Pins: the reference's own example input (02_lid_driven_cavity_coarse, a
data file) read by both readers (C++ and Python, written independently from
the grammar of src/femModel.cpp:3719-4213) with identical results; exact
round trips of generated files; the VTK text of a one-element model against
the format statements of src/femModel.cpp:1294-1530."""
import gzip
import os
import subprocess

import numpy as np
import pytest

from femorph_amd import meshgen, modelfile
from femorph_amd.build import DRIVER
from femorph_amd.mesh import BoundaryConditions, Mesh
from tests.vtkio import read_vtk_legacy

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _driver(*args, cwd):
    if not os.path.exists(DRIVER):
        pytest.fail("femx_explicit is not built: python -m femorph_amd.build")
    return subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=120)


def _dump(model_path, tmp_path):
    out = os.path.join(str(tmp_path), "dump.txt")
    r = _driver("--dump-model", out, model_path, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    d = dict(node=[], element=[], vel=[], dir=[], pres=[], ini=[], elvel=[], eldiff=[])
    for ln in open(out):
        t = ln.split()
        if t[0] in d:
            d[t[0]].append(t[1:])
        else:
            d[t[0]] = t[1:]
    return d


def _compare(d, m: modelfile.Model):
    """C++ reader dump vs the Python reader, exactly."""
    assert int(d["nodes"][0]) == len(m.coords)
    np.testing.assert_array_equal(np.array(d["node"], float).reshape(-1, 3), m.coords)
    assert int(d["elements"][0]) == len(m.conn)
    for row, ty, c in zip(d["element"], m.elem_types, m.conn):
        assert int(row[1]) == len(c) == modelfile.NODES_OF_TYPE[ty]
        assert [int(v) for v in row[2:]] == list(c)
    bc = m.bc
    assert len(d["vel"]) == len(bc.vel_node)
    for i, row in enumerate(d["vel"]):
        assert int(row[0]) == bc.vel_node[i] and int(row[1]) == bc.vel_dir[i]
        rows = np.array(row[3:], float).reshape(-1, 4)
        assert int(row[2]) == len(rows)
        np.testing.assert_array_equal(rows, bc.vel_rows[bc.vel_row_off[i]:bc.vel_row_off[i + 1]])
    np.testing.assert_array_equal(np.array([int(r[0]) for r in d["dir"]], np.int32), bc.dir_node)
    np.testing.assert_array_equal(np.array([r[1:] for r in d["dir"]], float).reshape(-1, 3), bc.dir_val)
    np.testing.assert_array_equal(np.array([int(r[0]) for r in d["pres"]], np.int32), bc.pres_node)
    np.testing.assert_array_equal(np.array([float(r[1]) for r in d["pres"]]), bc.pres_val)
    assert [float(v) for v in d["vmsprops"]] == m.vmsprops
    assert [float(v) for v in d["exopts"]] == m.exopts
    assert int(d["nodedof"][0]) == m.max_node_dofs
    assert d["nodedof"][1] == (m.fixed_dofs or "-")
    assert (float(d["timestep"][0]), int(d["timestep"][1]), int(d["timestep"][2])) == (
        m.time_step, m.total_steps, m.save_every)
    assert (float(d["dualstep"][0]), int(d["dualstep"][1])) == (m.dual_step, m.tot_dual_steps)
    assert float(d["exitcond"][0]) == m.exit_condition
    assert int(d["problem"][0]) == m.problem
    assert d["output"][0] == (m.output or "-") and d["logfile"][0] == (m.log_file or "-")
    assert [float(v) for v in d["timeintegration"]] == [m.alpha_m, m.alpha_f, m.gamma]
    assert [int(v) for v in d["stages"]] == m.stages
    assert int(d["prescribedvel"][0]) == m.prescribed_vel
    assert [(int(a), int(b), float(c)) for a, b, c in d["ini"]] == [tuple(r) for r in m.ini]
    assert [[float(v) for v in r] for r in d["elvel"]] == m.el_velocity
    assert [[float(v) for v in r] for r in d["eldiff"]] == m.el_diffusivity


def test_reference_example_model_both_readers(tmp_path):
    path = os.path.join(str(tmp_path), "model.txt")
    with gzip.open(os.path.join(GOLD, "ldc_coarse_model.txt.gz"), "rb") as f:
        open(path, "wb").write(f.read())
    m = modelfile.read_model_txt(path)
    # counts of the example's lines (grep of the file's keywords)
    assert len(m.coords) == 1118 and len(m.conn) == 4444 and set(m.elem_types) == {"TET4"}
    assert len(m.bc.dir_node) == 635 and int(m.bc.vel_row_off[-1]) == 298
    assert len(m.bc.vel_node) == 149  # two time rows (t = 0, 100) per lid node
    assert m.vmsprops == [1.06, 0.04, 2000.0, 4.0, 2.0] and (m.time_step, m.total_steps, m.save_every) == (0.001, 100, 10)
    assert m.problem == modelfile.PROBLEMS["EXPLICITVMS"]
    _compare(_dump(path, tmp_path), m)
    # every element is a valid, positively oriented tet (the VMS volume is signed)
    X = m.coords[np.array(m.conn)]
    det = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    assert (det != 0).all()


@pytest.mark.parametrize("name,counts", [
    # nodes, elements, NODEDIRBC, NODEVEL rows, velocity-BC nodes, NODEPRES, (dt, steps, save), (dual, tot dual)
    ("cylinder", (8595, 45754, 2565, 202, 101, 102, (0.01, 5000, 10), (1.0e-3, 0))),
    ("ldc_medium", (7309, 36468, 2392, 1022, 511, 0, (1.0e-3, 2, 1), (1.0e-6, 100))),
])
def test_reference_examples_both_readers(tmp_path, name, counts):
    """01_cylinder and 03_lid_driven_cavity_medium of the reference's
    examples: the counts of their keyword lines (grep of the files), and the
    C++ and Python readers agree exactly."""
    path = os.path.join(str(tmp_path), "model.txt")
    with gzip.open(os.path.join(GOLD, f"{name}_model.txt.gz"), "rb") as f:
        open(path, "wb").write(f.read())
    m = modelfile.read_model_txt(path)
    nn, ne, ndir, nvel_rows, nvel, npres, ts, ds = counts
    assert len(m.coords) == nn and len(m.conn) == ne and set(m.elem_types) == {"TET4"}
    assert len(m.bc.dir_node) == ndir and int(m.bc.vel_row_off[-1]) == nvel_rows and len(m.bc.vel_node) == nvel
    assert len(m.bc.pres_node) == npres
    assert (m.time_step, m.total_steps, m.save_every) == ts and (m.dual_step, m.tot_dual_steps) == ds
    assert m.vmsprops == [1.06, 0.04, 400.0, 4.0, 2.0]
    _compare(_dump(path, tmp_path), m)
    X = m.coords[np.array(m.conn)]
    det = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    assert (det > 0).all()


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_generated_model_round_trip(tmp_path, kind):
    if kind == "tet":
        mesh = meshgen.tet_cube(3, 0.15, 4)
        bc = meshgen.channel_bc(mesh)
        extra = dict(vmsprops=(1.0, 0.005, 1.0, 4.0, 2.0), max_node_dofs=4, tot_dual_steps=2, dual_step=0.003)
    else:
        mesh = meshgen.hex_cube(3, 0.2, 4)
        bc = meshgen.lid_cavity_bc(mesh)
        extra = dict(vmsprops=(1.0, 0.0025, 1.0, 4.0, 0.0), max_node_dofs=3, exopts=(0, 1, 1, 1, 1, 1),
                     fixed_dofs="Z", exit_condition=1.0e5, output="res", log_file="solver.log")
    path = os.path.join(str(tmp_path), "model.txt")
    modelfile.write_model_txt(path, mesh, bc, time_step=0.0125, total_steps=7, save_every=3, **extra)
    m = modelfile.read_model_txt(path)
    np.testing.assert_array_equal(m.coords, mesh.coords)
    np.testing.assert_array_equal(m.mesh().conn, mesh.conn)
    for f in ("vel_node", "vel_dir", "vel_row_off", "vel_rows", "dir_node", "dir_val", "pres_node", "pres_val"):
        np.testing.assert_array_equal(getattr(m.bc, f), getattr(bc, f), err_msg=f)
    _compare(_dump(path, tmp_path), m)


def test_grammar_details(tmp_path):
    """Case-insensitive keywords, ',' and repeated separators, 1-based ids,
    default NODEVEL direction 6, single-value NODEDIRBC, per-node time tables
    in line order, unknown keywords skipped."""
    text = "\n".join([
        "node 1 0 0 0", "NODE 2, 1.0, 0, 0", "Node 3  0,1 ,0", "NODE 4 0 0 1", "NODE 5 1 1 1",
        "element tet4 1 1 1 2 3 4", "ELEMENT TET4 2 1 2 3 4 5",
        "NODEVEL 5 1 0 0 0.5", "NODEVEL 2 0 1 0 0 1", "NODEVEL 5 2 0 0 1.5", "NODEVEL 2 0 3 0 2 1",
        "NODEDIRBC 3 0.25", "NODEDIRBC 4 1 2 3", "NODEPRES 1 -1.5",
        "INI 1 0 3.0", "COMMENT anything", "", "   ",
        "NODEDOF 4 xy", "TIMESTEP 1e-3 11 5", "DUALSTEP 2e-3 1", "VMSPROPS 1 2 3 4 5", "exitcond 7",
        "PROBLEM explicitvms",
    ]) + "\n"
    path = os.path.join(str(tmp_path), "model.txt")
    open(path, "w").write(text)
    m = modelfile.read_model_txt(path)
    assert m.coords.tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]]
    assert m.conn == [[0, 1, 2, 3], [1, 2, 3, 4]]
    assert m.bc.vel_node.tolist() == [1, 4]
    assert m.bc.vel_rows.tolist() == [[0, 1, 0, 0], [0, 3, 0, 2], [1, 0, 0, 0.5], [2, 0, 0, 1.5]]
    # direction codes are those of the first two NODEVEL lines (reference indexing)
    assert m.bc.vel_dir.tolist() == [6, 1]
    assert m.bc.dir_val.tolist() == [[0.25, 0, 0], [1, 2, 3]]
    assert m.fixed_dofs == "XY" and m.max_node_dofs == 4 and m.exit_condition == 7.0
    assert m.ini == [(0, 0, 3.0)]
    _compare(_dump(path, tmp_path), m)


def test_implicit_solver_keywords(tmp_path):
    """TIMEINTEGRATION, STAGES, PRESCRIBEDVEL, INI, ELVELS and ELDIFF
    (src/femModel.cpp:3892-3925, 4089-4133), written and read back by both
    readers."""
    mesh = meshgen.tet_cube(2, 0.1, 1)
    bc = meshgen.lid_cavity_bc(mesh)
    rng = np.random.default_rng(3)
    ev = rng.normal(size=(mesh.n_elems, 3))
    ed = rng.uniform(0, 0.1, (mesh.n_elems, 3))
    path = os.path.join(str(tmp_path), "model.txt")
    modelfile.write_model_txt(path, mesh, bc, max_node_dofs=5, time_integration=(0.8, 0.6, 0.7), stages=(0, 1),
                              prescribed_vel=2, ini=[(3, 4, 1.25), (0, 1, -0.5)], el_velocity=ev, el_diffusivity=ed)
    m = modelfile.read_model_txt(path)
    assert (m.alpha_m, m.alpha_f, m.gamma) == (0.8, 0.6, 0.7) and m.stages == [0, 1] and m.prescribed_vel == 2
    assert m.ini == [(3, 4, 1.25), (0, 1, -0.5)]
    np.testing.assert_array_equal(np.array(m.el_velocity), ev)
    np.testing.assert_array_equal(np.array(m.el_diffusivity), ed)
    _compare(_dump(path, tmp_path), m)


def test_invalid_fixed_dofs_is_an_error(tmp_path):
    path = os.path.join(str(tmp_path), "model.txt")
    open(path, "w").write("NODE 1 0 0 0\nNODEDOF 3 XYZ\n")
    with pytest.raises(modelfile.ModelFormatError):
        modelfile.read_model_txt(path)
    r = _driver("--dump-model", "d.txt", path, cwd=str(tmp_path))
    assert r.returncode != 0
    assert "ERROR: Invalid NODEDOF Option Format." in r.stdout and "Program Terminated." in r.stdout


def test_vtk_text_one_tet(tmp_path):
    path = os.path.join(str(tmp_path), "model.txt")
    open(path, "w").write("NODE 1 0 0 0\nNODE 2 1 0 0\nNODE 3 0 1 0\nNODE 4 0 0 1\nELEMENT TET4 1 1 1 2 3 4\n"
                          "NODEVEL 1 1 0 0 0\nNODEPRES 2 0\nNODEDIRBC 3 0 0 0\n")
    r = _driver("--write-vtk", "out.vtk", path, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout
    z3 = "0.000000e+00 0.000000e+00 0.000000e+00"
    expected = [
        "# vtk DataFile Version 2.0", "Model Exported from feMorph", "ASCII", "DATASET UNSTRUCTURED_GRID",
        "POINTS 4 float", z3, "1.000000e+00 0.000000e+00 0.000000e+00", "0.000000e+00 1.000000e+00 0.000000e+00",
        "0.000000e+00 0.000000e+00 1.000000e+00",
        "CELLS 1 5", "4 0 1 2 3 ", "CELL_TYPES 1", "10",
        "POINT_DATA 4", "VECTORS DXYZ double", z3, z3, z3, z3,
        "SCALARS dirBC double 1", "LOOKUP_TABLE default",
        "2.000000e+00", "3.000000e+00", "1.000000e+00", "0.000000e+00",
        "SCALARS neuBC double", "LOOKUP_TABLE default"] + ["0.000000e+00"] * 4 + [
        "CELL_DATA 1", "VECTORS normal double", z3,
        "SCALARS source double 1", "LOOKUP_TABLE default", "0.000000e+00", ""]
    assert open(os.path.join(str(tmp_path), "out.vtk")).read().split("\n") == expected


def test_vtk_hexa_mesh(tmp_path):
    mesh = meshgen.hex_cube(2, 0.1, 1)
    bc = meshgen.lid_cavity_bc(mesh)
    path = os.path.join(str(tmp_path), "model.txt")
    modelfile.write_model_txt(path, mesh, bc, max_node_dofs=3)
    r = _driver("--write-vtk", "out.vtk", path, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout
    v = read_vtk_legacy(os.path.join(str(tmp_path), "out.vtk"))
    assert (v["cell_types"] == 12).all() and v["cells"] == mesh.conn.tolist()
    np.testing.assert_allclose(v["points"], mesh.coords, rtol=6e-7, atol=0)
    flags = np.zeros(mesh.n_nodes)
    flags[bc.vel_node] = 2.0
    flags[bc.dir_node] = 1.0
    np.testing.assert_array_equal(v["point_data"]["dirBC"][:, 0], flags)
