"""The oracle against the reference's own golden vector: form_lumped_mass
(src/femSUPGExplicitFluidSolver.cpp:507-598) of example 06 (100 x 100 x 1
hexa8, rho = 10) vs examples/.../06_2D_hexa_lid_driven_cavity/lumped_mass.txt
(fixture made by tests/golden/make_lumped_mass_fixture.py)."""
import os

import numpy as np

from femorph_amd.mesh import Mesh
from oracle import oracle

FIX = os.path.join(os.path.dirname(__file__), "golden", "lumped_mass_06.npz")


def test_lumped_mass_matches_reference_file():
    d = np.load(FIX)
    m = Mesh(d["coords"], d["conn"])
    lm = oracle.lumped_mass(m, float(d["rho"]))
    ref = d["l_M"]
    # the file holds 11 significant digits ('%25.10e')
    assert np.max(np.abs(lm - ref) / np.abs(ref)) < 1e-10
    # per-dof column sums = rho * volume = 10 * 1 * 1 * 0.01
    for k in range(3):
        assert abs(lm[k::3].sum() - 0.1) < 1e-12
