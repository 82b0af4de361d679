"""The full-size workloads through checks that do not depend on size, each
finishing in seconds: the 96^3 x 6 VMS channel (5,308,416 tets) and the
128^3 SUPG cavity (2,097,152 hexa).

- VMS, plain FP64: one step against the oracle, <= 1e-12 relative L2 on the
  state and 1e-11 on the subgrid velocities (the oracle's threaded step,
  bitwise its serial one, keeps this to a few seconds);
- VMS, mixed: the same step within 1e-5 of the oracle, and five steps
  bitwise the same on two handles;
- SUPG, FP64 and mixed: a uniform velocity without boundary conditions is a
  fixed point of K1..K4 (its relative node velocities are exactly zero, so
  every gradient, strain and divergence is exactly zero); three steps are
  bitwise the same on two handles; mixed within 1e-5 of FP64 after three
  steps.
"""
import os

import numpy as np
import pytest

from femorph_amd import meshgen
from femorph_amd.mesh import BoundaryConditions
from oracle.oracle import VMSOracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _threads():
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, min(n, 16))


@pytest.fixture(scope="module")
def c3():
    m = meshgen.tet_cube(96, 0.15, 1)
    assert m.n_elems == 5308416 and m.n_nodes == 912673
    bc = meshgen.channel_bc(m, u_max=1.0)
    params = dict(rho=1.0, mu=1.0 / 200.0, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * meshgen.min_tet_h(m))
    ic = meshgen.random_velocity(m, 1e-3, 1)
    o = VMSOracle(m, bc, params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"], params["dt"])
    o.set_ic(ic)
    o.step(1, threads=_threads())
    ref = (o.get("sol_n"), o.get("sdus"), o.get("scalars")[0])
    del o
    return m, bc, params, ic, ref


def test_c3_fp64_step_matches_oracle(gpu, c3):
    m, bc, params, ic, (ref_sol, ref_sdus, ref_t) = c3
    dm = gpu.DeviceMesh(m)
    g = gpu.VMSSolver(dm, bc, precision=gpu.FEMX_FP64, **params)
    g.set_state(ic)
    g.step(1)
    sol, sdus, t = g.get_state()
    assert rel_l2(sol, ref_sol) <= 1e-12
    assert rel_l2(sdus, ref_sdus) <= 1e-11
    assert t == ref_t


def test_c3_mixed_step_and_determinism(gpu, c3):
    m, bc, params, ic, (ref_sol, _, _) = c3
    dm = gpu.DeviceMesh(m)
    outs = []
    for k in range(2):
        g = gpu.VMSSolver(dm, bc, precision=gpu.FEMX_MIXED, **params)
        g.set_state(ic)
        g.step(1)
        if k == 0:
            sol = g.get_state()[0]
            assert rel_l2(sol[:, :3], ref_sol[:, :3]) <= 1e-5
        g.step(4)
        outs.append(g.get_state())
        del g
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.fixture(scope="module")
def c2_mesh():
    from femorph_amd import femx
    assert femx.device_count() > 0, "no HIP device visible: GPU tests need an MI355X"
    m = meshgen.hex_cube(128, 0.2, 1)
    assert m.n_elems == 2097152 and m.n_nodes == 2146689
    return m, femx.DeviceMesh(m)


def _supg(gpu, dm, bc, precision):
    return gpu.SUPGSolver(dm, bc, 1.0, 1.0 / 400.0, 1.0, 4.0, 0.1 / 128, precision=precision)


@pytest.mark.parametrize("precision", ["FEMX_FP64", "FEMX_MIXED"])
def test_c2_uniform_velocity_is_a_fixed_point(gpu, c2_mesh, precision):
    m, dm = c2_mesh
    u0 = np.tile(np.array([0.3, -0.2, 0.1]), (m.n_nodes, 1))
    g = _supg(gpu, dm, BoundaryConditions(), getattr(gpu, precision))
    g.set_state(u0)
    assert g.step(3) == 3
    np.testing.assert_array_equal(g.get_state()["sol"], u0)


def test_c2_determinism_and_mixed_error(gpu, c2_mesh):
    m, dm = c2_mesh
    bc = meshgen.lid_cavity_bc(m)
    ic = meshgen.random_velocity(m, 1e-3, 2)[:, :3]
    outs = {}
    for name in ("FEMX_MIXED", "FEMX_MIXED", "FEMX_FP64"):
        g = _supg(gpu, dm, bc, getattr(gpu, name))
        g.set_state(ic)
        g.step(3)
        outs.setdefault(name, []).append(g.get_state())
        del g
    a, b = outs["FEMX_MIXED"]
    for key in ("sol", "pres", "lam"):
        assert np.array_equal(a[key], b[key]), key
    assert rel_l2(a["sol"], outs["FEMX_FP64"][0]["sol"]) <= 1e-5
