"""Unions of VMS handles (continuous batching of resident solves,
femx_vms_union_*): members of different sizes and Reynolds numbers join at
different times and leave when their own step counts are done; every
member's final state and subgrid velocities equal, bit for bit, those of the
same handle stepped on its own (same blocks, same sums, same order)."""
import numpy as np
import pytest

from femorph_amd import meshgen

pytestmark = pytest.mark.gpu


def _member(gpu, n, kind, jitter, seed, re):
    m = meshgen.tet_cube(n, jitter, seed)
    bc = meshgen.lid_cavity_bc(m) if kind == "cavity" else meshgen.channel_bc(m)
    p = dict(rho=1.0, mu=1.0 / re, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * meshgen.min_tet_h(m))
    ic = meshgen.random_velocity(m, 1e-3, seed + 11)
    return m, bc, p, ic


def _solver(gpu, case):
    m, bc, p, ic = case
    dm = gpu.DeviceMesh(m)
    g = gpu.VMSSolver(dm, bc, precision=gpu.FEMX_MIXED, **p)
    g.set_state(ic)
    return dm, g


def test_union_members_equal_their_own_runs(gpu):
    cases = [_member(gpu, 7, "cavity", 0.1, 1, 100.0), _member(gpu, 4, "channel", 0.15, 2, 200.0),
             _member(gpu, 11, "cavity", 0.0, 3, 400.0), _member(gpu, 3, "channel", 0.2, 4, 50.0),
             _member(gpu, 9, "channel", 0.1, 5, 300.0)]
    steps = [9, 14, 5, 11, 7]
    refs = []
    for case, n in zip(cases, steps):
        dm, g = _solver(gpu, case)
        g.step(n)
        sol, sdus, t = g.get_state()
        refs.append((sol, sdus, t))
        del g, dm
    u = gpu.VMSUnion()
    live = {}
    order = [0, 1, 2, 3, 4]
    joined = 0
    rounds = 0
    while joined < len(order) or live:
        # one new member per round: each runs its first step on its own
        if joined < len(order):
            i = order[joined]
            dm, g = _solver(gpu, cases[i])
            g.step(1)
            u.add(g, steps[i] - 1)
            live[i] = (dm, g)
            joined += 1
        print(f"round {rounds}: members {sorted(live)}", flush=True)
        u.step(2)
        rounds += 1
        for i in list(live):
            dm, g = live[i]
            if u.left(g) == 0:
                u.remove(g)
                sol, sdus, t = g.get_state()
                rsol, rsdus, rt = refs[i]
                assert t == rt, (i, t, rt)
                assert np.array_equal(sol, rsol), (i, np.abs(sol - rsol).max())
                assert np.array_equal(sdus, rsdus), (i, np.abs(sdus - rsdus).max())
                del live[i]
    assert rounds >= 5


def test_union_rejects_non_members(gpu):
    case = _member(gpu, 3, "cavity", 0.0, 1, 100.0)
    dm, g = _solver(gpu, case)
    u = gpu.VMSUnion()
    with pytest.raises(gpu.FemxError, match="first step"):
        u.add(g, 3)  # no step run on its own yet
    g.step(1)
    u.add(g, 3)
    with pytest.raises(gpu.FemxError, match="in a union"):
        g.step(1)
    with pytest.raises(gpu.FemxError, match="in a union"):
        g.get_state()
    u.step(3)
    u.remove(g)
    sol, _, _ = g.get_state()
    assert np.isfinite(sol).all()
    m, bc, p, ic = case
    g64 = gpu.VMSSolver(dm, bc, precision=gpu.FEMX_FP64, **p)
    g64.set_state(ic)
    g64.step(1)
    with pytest.raises(gpu.FemxError, match="FEMX_MIXED"):
        u.add(g64, 2)


def _supg_member(gpu, n, seed, re, exit_condition=1.0e5, ramp=False):
    m = meshgen.hex_cube(n, 0.2, seed)
    bc = meshgen.lid_cavity_bc(m)
    if ramp:
        # the lid speeds up linearly, 3 t, so max |u| passes 1 after t = 1/3
        rows = bc.vel_rows.reshape(-1, 2, 4).copy()
        rows[:, 0] = [0.0, 0.0, 0.0, 0.0]
        rows[:, 1] = [300.0, 0.0, 0.0, 100.0]
        bc.vel_rows = rows.reshape(-1, 4)
    return m, bc, dict(rho=1.0, mu=1.0 / re, alpha=1.0, cI=4.0, dt=0.1 / n, exit_condition=exit_condition)


def _supg_solver(gpu, case):
    m, bc, p = case
    dm = gpu.DeviceMesh(m)
    g = gpu.SUPGSolver(dm, bc, precision=gpu.FEMX_MIXED, **p)
    g.set_state(np.zeros((m.n_nodes, 3)))
    return dm, g


def test_supg_union_members_equal_their_own_runs(gpu):
    """SUPG unions: hexa8 members of different sizes and Reynolds numbers
    join at different times; every member's state, pressure, lambda and log
    row equal, bit for bit, those of the handle stepped on its own.  One
    member's exit condition fires inside the union: it stops at the same
    step as on its own."""
    # the 60^3 member has more log rows (element blocks and fix-up
    # workgroups) than one workgroup reduces on its own stream, where a
    # two-stage reduction makes its log row
    cases = [_supg_member(gpu, 7, 1, 100.0), _supg_member(gpu, 8, 2, 300.0), _supg_member(gpu, 10, 3, 50.0),
             _supg_member(gpu, 6, 4, 200.0, exit_condition=1.0, ramp=True), _supg_member(gpu, 9, 5, 400.0),
             _supg_member(gpu, 60, 6, 150.0)]
    steps = [9, 14, 5, 40, 11, 7]
    refs = []
    for case, n in zip(cases, steps):
        dm, g = _supg_solver(gpu, case)
        done = g.step(n)
        st = g.get_state()
        refs.append((done, st, g.get_log()))
        del g, dm
    assert refs[3][0] < steps[3], "the low exit condition must fire inside the run"
    u = gpu.SUPGUnion()
    live, joined, rounds = {}, 0, 0
    while joined < len(cases) or live:
        if joined < len(cases):
            i = joined
            dm, g = _supg_solver(gpu, cases[i])
            g.step_async(1)  # the first step on its own stream
            u.add(g, steps[i] - 1)
            live[i] = (dm, g)
            joined += 1
        print(f"supg round {rounds}: members {sorted(live)}", flush=True)
        u.step(3)
        rounds += 1
        for i in list(live):
            dm, g = live[i]
            if u.left(g) == 0:
                u.remove(g)
                done = g.finish()
                st = g.get_state()
                ref_done, ref, ref_log = refs[i]
                assert done == ref_done, (i, done, ref_done)
                for key in ("sol", "pres", "lam", "regime"):
                    np.testing.assert_array_equal(st[key], ref[key], err_msg=f"member {i} {key}")
                assert st["time"] == ref["time"]
                np.testing.assert_array_equal(g.get_log(), ref_log)
                del live[i]
    assert rounds > 3


def test_supg_union_rejects_non_members(gpu):
    """A SUPG union takes FEMX_MIXED hexa8 handles with B-hat, no fixed DOFs
    and no diagnostics, after their first step; a member is not stepped or
    read on its own until it leaves."""
    case = _supg_member(gpu, 6, 1, 100.0)
    dm, g = _supg_solver(gpu, case)
    u = gpu.SUPGUnion()
    with pytest.raises(gpu.FemxError, match="first step"):
        u.add(g, 3)
    g.step(1)
    u.add(g, 3)
    with pytest.raises(gpu.FemxError, match="in a union"):
        g.step(1)
    with pytest.raises(gpu.FemxError, match="in a union"):
        g.get_state()
    u.step(3)
    u.remove(g)
    assert g.finish() == 3  # the union steps count as its own enqueued steps
    assert np.isfinite(g.get_state()["sol"]).all()
    m, bc, p = case
    others = [gpu.SUPGSolver(dm, bc, precision=gpu.FEMX_FP64, **p),
              gpu.SUPGSolver(dm, bc, precision=gpu.FEMX_MIXED, fixed_dofs="Z", **p),
              gpu.SUPGSolver(dm, bc, precision=gpu.FEMX_MIXED, exopts=(0, 0, 1, 1, 1, 1), **p)]
    mt = meshgen.tet_cube(4, 0.1, 2)
    dmt = gpu.DeviceMesh(mt)
    others.append(gpu.SUPGSolver(dmt, meshgen.lid_cavity_bc(mt), precision=gpu.FEMX_MIXED,
                                 exopts=(0, 0, 1, 1, 1, 1), **p))
    gd = gpu.SUPGSolver(dm, bc, precision=gpu.FEMX_MIXED, **p)
    gd.set_diagnostics(True)
    others.append(gd)
    for h in others:
        h.set_state(np.zeros((h.N, 3)))
        h.step(1)
        assert not gpu.load().femx_supg_union_accepts(h._h)
        with pytest.raises(gpu.FemxError, match="union members"):
            u.add(h, 2)
