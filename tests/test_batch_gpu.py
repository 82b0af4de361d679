"""Batches of VMS tet4 meshes (femx_vms_batch_*): one launch per phase for
every mesh, steps replayed from a captured hipGraph.

A batch is the reference's time loop run for several models at once, so
each mesh must give exactly what its own single-mesh solve gives: the
tests require bitwise equality with separate femx_vms handles (same block
size), and the FP64 oracle bound of 1e-12 per step for one mesh of the
batch."""
import numpy as np
import pytest

from femorph_amd import meshgen
from oracle.oracle import VMSOracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _items(seed=0, precision=0, tot_dual=0, sizes=(3, 5, 4, 6, 2)):
    rng = np.random.default_rng(seed)
    items, ics = [], []
    for k, n in enumerate(sizes):
        m = meshgen.tet_cube(n, 0.15, seed + k)
        bc = meshgen.lid_cavity_bc(m)
        re = float(rng.uniform(50, 500))
        dt = 0.1 * meshgen.min_tet_h(m)
        items.append((m, bc, dict(rho=1.0, mu=1.0 / re, inv_eps=1.0, c1=4.0, c2=2.0, dt=dt,
                                  tot_dual_steps=tot_dual, precision=precision)))
        ics.append(meshgen.random_velocity(m, 1e-3, seed + 10 + k))
    return items, ics


def _singles(femx, items, ics, steps):
    out = []
    for (m, bc, p), ic in zip(items, ics):
        dm = femx.DeviceMesh(m)
        g = femx.VMSSolver(dm, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"],
                           tot_dual_steps=p["tot_dual_steps"], precision=p["precision"])
        g.set_state(ic)
        g.step(steps)
        sol, _, t = g.get_state()
        out.append((sol, t))
    return out


@pytest.mark.parametrize("precision,tot_dual", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0)])
def test_batch_equals_separate_solves(gpu, precision, tot_dual):
    items, ics = _items(1, precision, tot_dual)
    print("batch of", [m.n_elems for m, _, _ in items], "precision", precision, "dual", tot_dual, flush=True)
    b = gpu.VMSBatch(items, block_elems=256, xcd_order=1)
    info = b.info()
    assert info["meshes"] == 5 and info["elements"] == sum(m.n_elems for m, _, _ in items)
    b.set_state(ics)
    b.step(7)  # one eager step, then graph replays of two steps
    ref = _singles(gpu, items, ics, 7)
    for i, (sol_ref, t_ref) in enumerate(ref):
        sol, t = b.get_state(i)
        assert t == t_ref
        assert np.array_equal(sol, sol_ref), (i, rel_l2(sol, sol_ref))


def test_batch_graph_equals_eager(gpu):
    items, ics = _items(2)
    outs = []
    for graph in (True, False):
        b = gpu.VMSBatch(items, xcd_order=0)
        b.set_state(ics)
        b.step(5, graph=graph)
        b.step(4, graph=graph)
        outs.append([b.get_state(i)[0] for i in range(len(items))])
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


def test_batch_member_vs_oracle(gpu):
    items, ics = _items(3, sizes=(4, 3))
    b = gpu.VMSBatch(items)
    b.set_state(ics)
    m, bc, p = items[1]
    o = VMSOracle(m, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"])
    o.set_ic(ics[1])
    for step in range(6):
        b.step(1)
        o.step()
        sol, t = b.get_state(1)
        assert rel_l2(sol, o.get("sol_n")) <= 1e-12, step
        assert t == o.get("scalars")[0]


def test_batch_mixed_dual_member_vs_oracle(gpu):
    """Mixed precision with a second dual iteration per step, one member of
    a batch against the FP64 oracle (velocity <= 1e-5)."""
    items, ics = _items(5, precision=1, tot_dual=1, sizes=(5, 4, 6))
    b = gpu.VMSBatch(items)
    b.set_state(ics)
    m, bc, p = items[2]
    o = VMSOracle(m, bc, p["rho"], p["mu"], p["inv_eps"], p["c1"], p["c2"], p["dt"], tot_dual_steps=1)
    o.set_ic(ics[2])
    b.step(12)
    o.step(12)
    sol, t = b.get_state(2)
    assert t == o.get("scalars")[0]
    assert rel_l2(sol[:, :3], o.get("sol_n")[:, :3]) <= 1e-5


def test_batch_c4_shaped(gpu):
    """A C4-shaped batch at reduced count: 40 cavities, n log-uniform in
    [8, 20] (3,072 .. 48,000 tets each, uneven), Re uniform in [50, 500],
    mixed precision; every member bitwise equal to its own single-mesh solve."""
    rng = np.random.default_rng(21)
    ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(20), 40))).astype(int)
    items, ics, cache = [], [], {}
    for i, n in enumerate(ns):
        if n not in cache:
            m = meshgen.tet_cube(int(n))
            cache[n] = (m, meshgen.lid_cavity_bc(m), meshgen.min_tet_h(m))
        m, bc, h = cache[n]
        re = float(rng.uniform(50, 500))
        items.append((m, bc, dict(rho=1.0, mu=1.0 / re, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * h,
                                  tot_dual_steps=0, precision=1)))
        ics.append(meshgen.random_velocity(m, 1e-3, 100 + i))
    print("C4-shaped batch:", len(items), "meshes,", sum(m.n_elems for m, _, _ in items), "tets", flush=True)
    b = gpu.VMSBatch(items, block_elems=256, xcd_order=1)
    b.set_state(ics)
    b.step(12)
    ref = _singles(gpu, items, ics, 12)
    for i, (sol_ref, t_ref) in enumerate(ref):
        sol, t = b.get_state(i)
        assert t == t_ref
        assert np.array_equal(sol, sol_ref), (i, int(ns[i]), rel_l2(sol, sol_ref))


def test_autotune_picks_a_candidate(gpu):
    items, _ = _items(4, sizes=(6, 3, 5))
    be, xcd, ms = gpu.vms_batch_autotune(items, steps=4)
    print("autotune", be, xcd, ms, flush=True)
    assert be in (256, 224, 192) and xcd in (0, 1)
    assert (ms > 0).all() and ms.min() == ms[[256, 224, 192].index(be) * 2 + (1 - xcd)]


def test_one_mesh_batch_time_dependent_bc(gpu):
    """A batch of one mesh reads its time from the device table like any
    batch: with a lid velocity ramped in time, it must equal the
    single-mesh solver (which passes the host time) bitwise."""
    from femorph_amd.mesh import BoundaryConditions
    m = meshgen.tet_cube(4, 0.1, 7)
    base = meshgen.lid_cavity_bc(m)
    dt = 0.1 * meshgen.min_tet_h(m)
    n_lid = len(base.vel_node)
    rows = np.zeros((2 * n_lid, 4))
    rows[1::2] = [1.0, 0.0, 0.0, 10 * dt]  # u = t / (10 dt) up to t = 10 dt
    bc = BoundaryConditions(vel_node=base.vel_node, vel_dir=base.vel_dir,
                            vel_row_off=np.arange(0, 2 * n_lid + 1, 2, dtype=np.int32), vel_rows=rows,
                            dir_node=base.dir_node, dir_val=base.dir_val)
    p = dict(rho=1.0, mu=0.01, inv_eps=1.0, c1=4.0, c2=2.0, dt=dt, tot_dual_steps=0, precision=0)
    ic = meshgen.random_velocity(m, 1e-3, 5)
    b = gpu.VMSBatch([(m, bc, p)])
    b.set_state([ic])
    b.step(6)
    sol_b, t_b = b.get_state(0)
    ref = _singles(gpu, [(m, bc, p)], [ic], 6)[0]
    assert t_b == ref[1] and np.array_equal(sol_b, ref[0])
    lid = sol_b[np.setdiff1d(bc.vel_node, bc.dir_node), 0]  # lid edges are walls (walls win)
    assert len(lid) > 0 and np.allclose(lid, t_b / (10 * dt))  # the ramp at t = 6 dt
