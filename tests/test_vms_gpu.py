"""VMS tet4 on the GPU vs the FP64 oracle (restatement of
src/femVMSExplicitFluidSolver.cpp), through the C ABI.

Tolerances: plain FP64 <= 1e-12 relative L2 per step (state u v w p and the
subgrid velocity); mixed FP32-element / FP64-state: stated per test."""
import numpy as np
import pytest

from femorph_amd import meshgen
from oracle.oracle import VMSOracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _case(kind, n, jitter=0.0, seed=0):
    if kind == "cavity":
        m = meshgen.tet_cube(n, jitter, seed)
        bc = meshgen.lid_cavity_bc(m)
        mu = 0.01
    else:
        m = meshgen.tet_cube(n, jitter, seed)
        bc = meshgen.channel_bc(m)
        mu = 1.0 / 200.0
    dt = 0.1 * meshgen.min_tet_h(m)
    params = dict(rho=1.0, mu=mu, inv_eps=1.0, c1=4.0, c2=2.0, dt=dt)
    ic = meshgen.random_velocity(m, 1e-3, seed + 7)
    return m, bc, params, ic


def _pair(femx, m, bc, params, ic, precision=0, tot_dual_steps=0, sdus=None, block_elems=0, order=0):
    o = VMSOracle(m, bc, params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"], params["dt"],
                  tot_dual_steps=tot_dual_steps)
    o.set_ic(ic)
    if sdus is not None:
        o.set_sdus(sdus)
    dm = femx.DeviceMesh(m, block_elems, order)
    g = femx.VMSSolver(dm, bc, params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"],
                       params["dt"], tot_dual_steps=tot_dual_steps, precision=precision)
    g.set_state(ic, sdus)
    return o, g, dm


def test_mass_and_initial_state(gpu):
    m, bc, params, ic = _case("cavity", 6)
    o, g, _ = _pair(gpu, m, bc, params, ic)
    assert rel_l2(g.get_mass(), o.get("nd_mass")) <= 1e-15
    sol, sdus, t = g.get_state()
    assert np.array_equal(sol, o.get("sol_n"))
    assert t == 0.0 and not sdus.any()


@pytest.mark.parametrize("block_elems,order", [(0, 1), (64, 0), (100, 1)])
def test_block_choices_keep_parity(gpu, block_elems, order):
    """Blocks in the mesh's element order, and smaller blocks: per-step FP64
    parity with the oracle and mixed precision within 1e-5."""
    m, bc, params, ic = _case("channel", 6, 0.15, seed=4)
    st = gpu.block_stats(m, block_elems or 256, order)
    assert st["max_elems"] <= (block_elems or 256) and st["blocks"] >= 5
    o, g, _ = _pair(gpu, m, bc, params, ic, block_elems=block_elems, order=order)
    for step in range(6):
        o.step()
        g.step()
        sol, sdus, _ = g.get_state()
        assert rel_l2(sol, o.get("sol_n")) <= TOL64, (step, rel_l2(sol, o.get("sol_n")))
        assert rel_l2(sdus, o.get("sdus")) <= 1e-11
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=1, block_elems=block_elems, order=order)
    for step in range(6):
        o.step()
        g.step()
    assert rel_l2(g.get_state()[0], o.get("sol_n")) <= 1e-5


@pytest.mark.parametrize("kind,n,jitter", [("cavity", 6, 0.0), ("channel", 6, 0.15), ("cavity", 5, 0.2)])
def test_fp64_per_step_parity(gpu, kind, n, jitter):
    m, bc, params, ic = _case(kind, n, jitter, seed=3)
    o, g, _ = _pair(gpu, m, bc, params, ic)
    for step in range(25):
        o.step()
        g.step()
        sol, sdus, t = g.get_state()
        assert rel_l2(sol, o.get("sol_n")) <= TOL64, step
        assert rel_l2(sdus, o.get("sdus")) <= 1e-11, step
        assert t == o.get("scalars")[0]


@pytest.mark.parametrize("precision,tol", [(0, TOL64), (1, 1e-5), (2, 1e-5)])
def test_nonzero_pressure_and_offset_velocity(gpu, precision, tol):
    """Every node starts with a non-integer pressure, so the pressure
    gradient the projection takes from global nodes 0..3 (:489) is nonzero,
    and the velocity carries a non-integer mean, so the block reference
    values of FP32 element math are not small integers."""
    m, bc, params, ic = _case("channel", 6, 0.15, seed=9)
    rng = np.random.default_rng(9)
    ic[:, 3] = rng.uniform(0.2, 0.8, m.n_nodes)
    ic[:, :3] += np.array([0.37, -0.21, 0.13])
    assert np.ptp(ic[:4, 3]) > 0.01
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=precision)
    for step in range(6):
        o.step()
        g.step()
    sol, _, _ = g.get_state()
    assert rel_l2(sol, o.get("sol_n")) <= tol, rel_l2(sol, o.get("sol_n"))


def test_fp64_fused_run_matches(gpu):
    """Continuous stepping (projection fused into the next residual launch)."""
    m, bc, params, ic = _case("channel", 6, 0.15, seed=5)
    o, g, _ = _pair(gpu, m, bc, params, ic)
    o.step(40)
    g.step(40)
    sol, sdus, _ = g.get_state()
    assert rel_l2(sol, o.get("sol_n")) <= TOL64
    assert rel_l2(sdus, o.get("sdus")) <= 1e-11
    mx, cfl = g.get_cfl()
    assert abs(cfl - o.get("scalars")[4]) <= 1e-12 * max(1.0, cfl)


@pytest.mark.parametrize("precision,tol,tol_sd", [(0, TOL64, 1e-11), (1, 1e-5, 1e-4), (2, 1e-5, 1e-4)])
def test_dual_iterations(gpu, precision, tol, tol_sd):
    """tot_dual_steps = 2: acceleration term and pre-BC base of later dual
    iterations, FP64 and both mixed modes (the FP32 dd = dr + (uq - up) path of
    the accumulating kernel).  On 8^3 x 6 tets the lid and wall nodes fall on
    many block boundaries, so first-slot base states, pre-BC records and the
    scaled partials all take part."""
    m, bc, params, ic = _case("cavity", 8, 0.1, seed=2)
    st = gpu.block_stats(m, 256)
    assert st["blocks"] >= 12 and st["shared_nodes"] > 200
    params["dt"] *= 0.5
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=precision, tot_dual_steps=2)
    for step in range(10):
        o.step()
        g.step()
        sol, sdus, _ = g.get_state()
        assert rel_l2(sol, o.get("sol_n")) <= tol, (step, rel_l2(sol, o.get("sol_n")))
        assert rel_l2(sdus, o.get("sdus")) <= tol_sd, (step, rel_l2(sdus, o.get("sdus")))


def test_fp64_nonzero_subscale_ic(gpu):
    m, bc, params, ic = _case("channel", 5, 0.1, seed=4)
    rng = np.random.default_rng(11)
    sd = rng.uniform(-1e-3, 1e-3, size=(12, m.n_elems))
    o, g, _ = _pair(gpu, m, bc, params, ic, sdus=sd)
    o.step(5)
    g.step(5)
    sol, sdus, _ = g.get_state()
    assert rel_l2(sol, o.get("sol_n")) <= TOL64
    assert rel_l2(sdus, o.get("sdus")) <= 1e-11


def test_fp64_deterministic(gpu):
    m, bc, params, ic = _case("channel", 7, 0.15, seed=8)
    dm = gpu.DeviceMesh(m)
    outs = []
    for _ in range(2):
        g = gpu.VMSSolver(dm, bc, **params)
        g.set_state(ic)
        g.step(15)
        outs.append(g.get_state())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_c1_cavity_100_steps(gpu):
    """C1: 20^3 x 6 = 48,000 tets, Re 100 lid cavity, 100 time steps (99 run,
    like the reference's loopTime = 1 .. totalSteps-1), per-step <= 1e-12."""
    m, bc, params, ic = _case("cavity", 20)
    assert m.n_elems == 48000 and m.n_nodes == 9261
    o, g, _ = _pair(gpu, m, bc, params, ic)
    worst = 0.0
    for step in range(99):
        o.step()
        g.step()
        if step % 11 == 0 or step == 98:
            sol, _, _ = g.get_state()
            worst = max(worst, rel_l2(sol, o.get("sol_n")))
    assert worst <= TOL64, worst


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_precision_error(gpu, precision):
    """FP32 element / FP64 state vs the FP64 oracle after 200 steps:
    FEMX_MIXED (1) and the opt-in FEMX_MIXED_F32SUM (2)."""
    m, bc, params, ic = _case("channel", 8, 0.15, seed=6)
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=precision)
    o.step(200)
    g.step(200)
    sol, _, _ = g.get_state()
    ref = o.get("sol_n")
    err_u = rel_l2(sol[:, :3], ref[:, :3])
    assert err_u <= 1e-5, err_u


@pytest.mark.parametrize("tot_dual", [0, 2])
def test_element_diagnostics(gpu, tot_dual):
    """res_v / res_p (centroid residuals of the partially assembled rhs,
    ascending element order), subgrid_u / subgrid_p and the printed norms
    and CFL of the last dual iteration vs the oracle, which assembles
    element by element like the reference.  Tolerance 1e-11: the running
    sums cancel, so their rounding is relative to the contributions."""
    m, bc, params, ic = _case("channel", 5, 0.15, seed=5)
    sdus = np.random.default_rng(3).uniform(-1e-3, 1e-3, (12, m.n_elems))
    o, g, _ = _pair(gpu, m, bc, params, ic, tot_dual_steps=tot_dual, sdus=sdus)
    g.set_diagnostics(True)
    for step in range(4):
        o.step()
        g.step()
        d = g.get_diagnostics()
        errs = dict(res_v=rel_l2(d["res_v"], o.get("el_u_res")), res_p=rel_l2(d["res_p"], o.get("el_p_res")),
                    subgrid_u=rel_l2(d["subgrid_u"], o.get("el_u_sdus")))
        print(step, errs, flush=True)
        assert max(errs.values()) <= 1e-11, errs
        assert not d["subgrid_p"].any()
        s = o.get("scalars")
        last = d["iter_norms"][-1]
        assert d["iter_norms"].shape == (tot_dual + 1, 3)
        np.testing.assert_allclose(last, s[2:5], rtol=1e-11, atol=0)
    # switching the diagnostics off leaves the state path untouched
    g.set_diagnostics(False)
    o.step()
    g.step()
    sol, _, _ = g.get_state()
    assert rel_l2(sol, o.get("sol_n")) <= TOL64


def test_element_diagnostics_mixed(gpu):
    """The same diagnostics with FP32 element math (the drop-in's --mixed
    solve): res_v / res_p and the printed norms within 1e-6 of the oracle.
    subgrid_u is bounded against the scale of the subgrid velocities: once
    the orthogonal projection has acted, its Gauss-point mean is zero up to
    rounding (about 1e-17 in FP64, FP32 rounding here)."""
    m, bc, params, ic = _case("channel", 5, 0.15, seed=5)
    sdus = np.random.default_rng(3).uniform(-1e-3, 1e-3, (12, m.n_elems))
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=1, sdus=sdus)
    g.set_diagnostics(True)
    for step in range(4):
        o.step()
        g.step()
        d = g.get_diagnostics()
        assert rel_l2(d["res_v"], o.get("el_u_res")) <= 1e-6, step
        assert rel_l2(d["res_p"], o.get("el_p_res")) <= 1e-6, step
        scale = np.abs(o.get("sdus")).max()
        assert np.abs(d["subgrid_u"] - o.get("el_u_sdus")).max() <= 1e-6 * scale, step
        np.testing.assert_allclose(d["iter_norms"][-1], o.get("scalars")[2:5], rtol=1e-6, atol=0)


def test_norms_ring_without_fields(gpu):
    """Diagnostics mode 2 (norms only): the per-iteration norms of 70 steps
    stay on the device and femx_vms_get_norms returns the last 64, oldest
    first, equal to the oracle's printed values of each step; a mode 1 step
    after them still yields the element fields."""
    m, bc, params, ic = _case("channel", 4, 0.15, seed=6)
    o, g, _ = _pair(gpu, m, bc, params, ic, tot_dual_steps=1)
    g.set_diagnostics(2)
    ref = []
    for step in range(70):
        o.step()
        g.step()
        ref.append(o.get("scalars")[2:5].copy())
    got = g.get_norms(64)
    assert got.shape == (64, 2, 3)
    np.testing.assert_allclose(got[:, -1, :], np.array(ref[6:]), rtol=1e-11, atol=0)
    with pytest.raises(gpu.FemxError, match="at most the last 64"):
        g.get_norms(65)
    g.set_diagnostics(1)
    o.step()
    g.step()
    d = g.get_diagnostics()
    assert rel_l2(d["res_v"], o.get("el_u_res")) <= 1e-11
    np.testing.assert_allclose(d["iter_norms"][-1], o.get("scalars")[2:5], rtol=1e-11, atol=0)
    np.testing.assert_allclose(g.get_norms(1)[0], d["iter_norms"], rtol=0, atol=0)


@pytest.mark.parametrize("precision,tol", [(0, TOL64), (1, 1e-5)])
def test_high_valence_node(gpu, precision, tol):
    """A node with 512 incident tets: the block builder splits its corners
    over blocks of at most 128 corners of one node; per-step parity with the
    oracle, no BCs."""
    m = meshgen.tet_fan(3)
    assert m.n_elems == 512 and np.bincount(m.conn.ravel())[0] == 512
    assert (meshgen._tet_det(m) > 0).all()
    s = gpu.block_stats(m, 256)
    assert s["blocks"] >= 4 and s["shared_nodes"] >= 1
    bc = meshgen.BoundaryConditions()
    dt = 0.1 * meshgen.min_tet_h(m)
    params = dict(rho=1.0, mu=0.01, inv_eps=1.0, c1=4.0, c2=2.0, dt=dt)
    ic = meshgen.random_velocity(m, 1e-3, 5)
    ic[:, 3] = np.random.default_rng(5).uniform(0.2, 0.8, m.n_nodes)
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=precision)
    for step in range(8):
        o.step()
        g.step()
        sol, _, _ = g.get_state()
        assert rel_l2(sol, o.get("sol_n")) <= tol, (step, rel_l2(sol, o.get("sol_n")))


@pytest.mark.parametrize("precision,tol", [(0, TOL64), (1, 1e-5), (2, 1e-5)])
def test_node_of_no_element(gpu, precision, tol):
    """A node that no element uses: the update divides its zero residual by
    its zero mass, as the reference does, so its state turns NaN; every
    other node keeps parity with the oracle."""
    m0, bc, params, _ = _case("channel", 4, 0.1, seed=3)
    m = meshgen.Mesh(np.vstack([m0.coords, [[5.0, 5.0, 5.0]]]), m0.conn)
    ic = meshgen.random_velocity(m, 1e-3, 3)
    o, g, _ = _pair(gpu, m, bc, params, ic, precision=precision)
    for step in range(4):
        o.step()
        g.step()
        sol, _, _ = g.get_state()
        ref = o.get("sol_n")
        assert np.isnan(ref[-1]).all() and np.isnan(sol[-1]).all(), (step, sol[-1])
        assert rel_l2(sol[:-1], ref[:-1]) <= tol, (step, rel_l2(sol[:-1], ref[:-1]))


def test_duplicate_velocity_node_rejected(gpu):
    """The reference walks a compacted (sorted, unique) velocity-node list,
    one time table per node (src/femModel.cpp:4558-4576, :4581-4628); a C-ABI
    caller that lists a node twice gets an argument error, for the VMS and
    the SUPG solver alike."""
    m, bc, params, _ = _case("cavity", 3)
    k = int(bc.vel_node[0])
    rows = bc.vel_rows[bc.vel_row_off[0]:bc.vel_row_off[1]]
    dup = meshgen.BoundaryConditions(
        vel_node=np.r_[bc.vel_node, k], vel_dir=np.r_[bc.vel_dir[:len(bc.vel_node)], 1],
        vel_row_off=np.r_[bc.vel_row_off, bc.vel_row_off[-1] + len(rows)], vel_rows=np.vstack([bc.vel_rows, rows]),
        dir_node=bc.dir_node, dir_val=bc.dir_val, pres_node=bc.pres_node, pres_val=bc.pres_val)
    dm = gpu.DeviceMesh(m)
    with pytest.raises(gpu.FemxError, match="listed twice"):
        gpu.VMSSolver(dm, dup, params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"], params["dt"])
    with pytest.raises(gpu.FemxError, match="listed twice"):
        gpu.SUPGSolver(dm, dup, rho=1.0, mu=0.01, alpha=1.0, cI=4.0, dt=params["dt"], exopts=(0, 0, 1, 1, 1, 1))
    # the unmodified tables still build
    gpu.VMSSolver(dm, bc, params["rho"], params["mu"], params["inv_eps"], params["c1"], params["c2"], params["dt"])
