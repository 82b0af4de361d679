"""ctypes front end of the FP64 oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (built from femx_oracle.c by oracle/Makefile).
Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OxMesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_elems", C.c_int), ("npe", C.c_int), ("coords", _dp), ("conn", _ip)]


class OxBC(C.Structure):
    _fields_ = [("n_vel", C.c_int), ("vel_node", _ip), ("vel_dir", _ip), ("vel_row_off", _ip), ("vel_rows", _dp),
                ("n_dir", C.c_int), ("dir_node", _ip), ("dir_val", _dp),
                ("n_pres", C.c_int), ("pres_node", _ip), ("pres_val", _dp)]


class OxVMSParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("inv_eps", C.c_double), ("c1", C.c_double),
                ("c2", C.c_double), ("dt", C.c_double), ("dual_step", C.c_double), ("tot_dual_steps", C.c_int)]


class OxSUPGParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("alpha", C.c_double), ("cI", C.c_double),
                ("dt", C.c_double), ("use_euler", C.c_int), ("use_bhat", C.c_int), ("k1", C.c_int),
                ("k2", C.c_int), ("k3", C.c_int), ("k4", C.c_int), ("fixed_dofs", C.c_int),
                ("exit_condition", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.environ.get("FEMX_ORACLE_LIB")  # a mutant of the restatement (mutation tests)
        if not path:
            path = os.path.join(_HERE, "liboracle.so")
            if not os.path.exists(path):
                build()
        L = C.CDLL(path)
        L.ox_vms_create.restype = C.c_void_p
        L.ox_vms_create.argtypes = [C.POINTER(OxMesh), C.POINTER(OxBC), C.POINTER(OxVMSParams)]
        L.ox_vms_destroy.argtypes = [C.c_void_p]
        L.ox_vms_set_ic.argtypes = [C.c_void_p, _dp]
        L.ox_vms_set_sdus.argtypes = [C.c_void_p, _dp]
        L.ox_vms_step.argtypes = [C.c_void_p]
        L.ox_vms_step_threads.argtypes = [C.c_void_p, C.c_int]
        L.ox_vms_get.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ox_vms_element_residual.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.ox_supg_create.restype = C.c_void_p
        L.ox_supg_create.argtypes = [C.POINTER(OxMesh), C.POINTER(OxBC), C.POINTER(OxSUPGParams)]
        L.ox_supg_destroy.argtypes = [C.c_void_p]
        L.ox_supg_set_ic.argtypes = [C.c_void_p, _dp]
        L.ox_supg_step.argtypes = [C.c_void_p]
        L.ox_supg_get.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ox_supg_set_elem_range.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ox_supg_lumped_mass.argtypes = [C.POINTER(OxMesh), C.c_double, _dp]
        L.ox_rule_second_order.argtypes = [C.c_int, _dp, _dp]
        L.ox_interp_table.argtypes = [C.c_double, _dp, C.c_int, _dp]
        L.ox_shape.argtypes = [C.c_int, _dp, _dp]
        L.ox_global_deriv.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.ox_geom_matrix.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ox_ins_pattern.restype = C.c_long
        L.ox_ins_pattern.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip]
        L.ox_ins_prescribe.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp]
        L.ox_ins_element.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp] + [C.c_double] * 4 + [_dp, _dp]
        L.ox_ins_assemble.argtypes = ([C.POINTER(OxMesh), _ip, _ip, _dp, _dp, _dp, _dp] + [C.c_double] * 4 +
                                      [C.c_int, _ip, _dp, _dp, _dp])
        L.ox_gmres.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int, _ip, _dp]
        L.ox_ins_update.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        _LIB = L
    return _LIB


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class _Keep:
    """Holds numpy arrays alive while C structs point into them."""

    def __init__(self, mesh, bc):
        self.coords = np.ascontiguousarray(mesh.coords, np.float64)
        self.conn = np.ascontiguousarray(mesh.conn, np.int32)
        self.m = OxMesh(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(self.coords), _p(self.conn, _ip))
        self.arrs = [np.ascontiguousarray(x) for x in (bc.vel_node, bc.vel_dir, bc.vel_row_off, bc.vel_rows,
                                                        bc.dir_node, bc.dir_val, bc.pres_node, bc.pres_val)]
        a = self.arrs
        self.b = OxBC(bc.n_vel, _p(a[0], _ip), _p(a[1], _ip), _p(a[2], _ip), _p(a[3]),
                      len(bc.dir_node), _p(a[4], _ip), _p(a[5]), len(bc.pres_node), _p(a[6], _ip), _p(a[7]))


class VMSOracle:
    """FP64 restatement of femVMSExplicitFluidSolver::solve, step by step."""

    GET = dict(sol_n=(0, "N4"), sol_tau=(1, "N4"), sdus=(2, "E12"), sgps=(3, "E4"), nd_mass=(4, "N"),
               volumes=(5, "E"), char_length=(6, "E"), DNs=(7, "E12"), elmTaus=(8, "E2"), el_u_res=(9, "E3"),
               el_p_res=(10, "E"), el_u_sdus=(11, "E3"), el_p_sdus=(12, "E"), scalars=(13, "S6"))

    def __init__(self, mesh, bc, rho, mu, inv_eps, c1, c2, dt, dual_step=None, tot_dual_steps=0):
        L = lib()
        self.N, self.E = mesh.n_nodes, mesh.n_elems
        k = _Keep(mesh, bc)
        p = OxVMSParams(rho, mu, inv_eps, c1, c2, dt, dt if dual_step is None else dual_step, tot_dual_steps)
        self._h = L.ox_vms_create(C.byref(k.m), C.byref(k.b), C.byref(p))
        if not self._h:
            raise ValueError("ox_vms_create failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ox_vms_destroy(self._h)
            self._h = None

    def set_ic(self, sol):
        sol = np.ascontiguousarray(sol, np.float64).reshape(self.N, 4)
        if lib().ox_vms_set_ic(self._h, _p(sol)) != 0:
            raise ValueError("BC table error")

    def set_sdus(self, sdus):
        sdus = np.ascontiguousarray(sdus, np.float64).reshape(12 * self.E)
        lib().ox_vms_set_sdus(self._h, _p(sdus))

    def step(self, n=1, threads=1):
        """n steps; threads > 1 runs the element sweeps on that many threads
        (ox_vms_step_threads: bitwise the same results)."""
        for _ in range(n):
            r = lib().ox_vms_step(self._h) if threads <= 1 else lib().ox_vms_step_threads(self._h, int(threads))
            if r != 0:
                raise ValueError("BC table error" if r == -1 else "out of host memory")

    def get(self, name):
        which, shape = self.GET[name]
        dims = {"N4": (self.N, 4), "E12": (12, self.E), "E4": (4, self.E), "N": (self.N,), "E": (self.E,),
                "E2": (2, self.E), "E3": (self.E, 3), "S6": (6,)}[shape]
        out = np.zeros(dims)
        lib().ox_vms_get(self._h, which, _p(out))
        return out

    def element_residual(self, e, cur, prev):
        cur = np.ascontiguousarray(cur, np.float64)
        prev = np.ascontiguousarray(prev, np.float64)
        out = np.zeros(16)
        lib().ox_vms_element_residual(self._h, e, _p(cur), _p(prev), _p(out))
        return out.reshape(4, 4)


_FIXED = {"NONE": 0, "X": 1, "Y": 2, "Z": 3, "XY": 4, "YZ": 5, "XZ": 6}


class SUPGOracle:
    """FP64 restatement of femSUPGExplicitFluidSolver::solve, step by step."""

    GET = dict(sol_n=(0, "N3"), sol_p=(1, "N3"), l_M=(2, "N3"), el_lam=(3, "E"), el_regime=(4, "E"),
               pres_n=(5, "E"), k_sol_p=(6, "E24"), k_sol_pp=(7, "E24"), log=(8, "S19"), scalars=(9, "S2"),
               l_cfl_step=(10, "S1"))

    def __init__(self, mesh, bc, rho, mu, alpha, cI, dt, exopts=(0, 1, 1, 1, 1, 1), fixed_dofs="NONE",
                 exit_condition=1.0e5):
        L = lib()
        self.N, self.E = mesh.n_nodes, mesh.n_elems
        k = _Keep(mesh, bc)
        e = [int(bool(x)) for x in exopts]
        p = OxSUPGParams(rho, mu, alpha, cI, dt, e[0], e[1], e[2], e[3], e[4], e[5], _FIXED[fixed_dofs],
                         exit_condition)
        self._h = L.ox_supg_create(C.byref(k.m), C.byref(k.b), C.byref(p))
        if not self._h:
            raise ValueError("ox_supg_create failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ox_supg_destroy(self._h)
            self._h = None

    def set_ic(self, sol):
        sol = np.ascontiguousarray(np.asarray(sol, np.float64)[:, :3]).reshape(self.N, 3)
        if lib().ox_supg_set_ic(self._h, _p(sol)) != 0:
            raise ValueError("BC table error")

    def set_elem_range(self, e0, e1):
        lib().ox_supg_set_elem_range(self._h, e0, e1)

    def step(self, n=1):
        """Returns 1 if the exit condition fired on the last step."""
        r = 0
        for _ in range(n):
            r = lib().ox_supg_step(self._h)
            if r < 0:
                raise ValueError("BC / option error")
            if r == 1:
                return 1
        return r

    def get(self, name):
        which, shape = self.GET[name]
        dims = {"N3": (self.N, 3), "E": (self.E,), "E24": (24, self.E), "S19": (19,), "S2": (2,), "S1": (1,)}[shape]
        out = np.zeros(dims)
        lib().ox_supg_get(self._h, which, _p(out))
        return out


def lumped_mass(mesh, rho):
    k = _Keep(mesh, __import__("femorph_amd.mesh", fromlist=["BoundaryConditions"]).BoundaryConditions())
    out = np.zeros(3 * mesh.n_nodes)
    lib().ox_supg_lumped_mass(C.byref(k.m), rho, _p(out))
    return out


def rule_second_order(npe):
    c = np.zeros(32)
    w = np.zeros(8)
    n = lib().ox_rule_second_order(npe, _p(c), _p(w))
    return c.reshape(8, 4)[:n], w[:n]


def ffd_morph(min_point, n, S, T, U, disp_node, disp_val, coords):
    """femPointGrid morph of points (box coordinates, Bernstein sum)."""
    L = lib()
    c = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    out = np.zeros_like(c)
    dn = np.ascontiguousarray(disp_node, np.int32)
    dv = np.ascontiguousarray(disp_val, np.float64).reshape(-1, 3)
    v = [np.ascontiguousarray(x, np.float64) for x in (min_point, S, T, U)]
    nn = np.ascontiguousarray(n, np.int32)
    r = L.ox_ffd_morph(_p(v[0]), _p(nn, _ip), _p(v[1]), _p(v[2]), _p(v[3]), len(dn), _p(dn, _ip), _p(dv), len(c),
                       _p(c), _p(out))
    if r != 0:
        raise ValueError("control point index out of range")
    return out


def interp_table(t, rows):
    rows = np.ascontiguousarray(rows, np.float64).reshape(-1, 4)
    out = np.zeros(3)
    r = lib().ox_interp_table(t, _p(rows), len(rows), _p(out))
    if r != 0:
        raise ValueError("time precedes table")
    return out


def shape(npe, xi):
    xi = np.ascontiguousarray(xi, np.float64)
    N = np.zeros(npe)
    lib().ox_shape(npe, _p(xi), _p(N))
    return N


def global_deriv(npe, xel, xi):
    xel = np.ascontiguousarray(xel, np.float64).reshape(npe, 3)
    xi = np.ascontiguousarray(xi, np.float64)
    det = np.zeros(1)
    DN = np.zeros((npe, 3))
    lib().ox_global_deriv(npe, _p(xel), _p(xi), _p(det), _p(DN))
    return float(det[0]), DN


def geom_matrix(npe, xel, xi):
    xel = np.ascontiguousarray(xel, np.float64).reshape(npe, 3)
    xi = np.ascontiguousarray(xi, np.float64)
    G = np.zeros(9)
    lib().ox_geom_matrix(npe, _p(xel), _p(xi), _p(G))
    return G.reshape(3, 3)


def ref_rules_lib():
    """The reference's own femIntegrationRule (oracle/_ref), or None."""
    path = os.path.join(_HERE, "_ref", "libref_rules.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_rule.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
    return L


def ins_pattern(mesh):
    """CSR pattern of femSparseMatrix (row_ptr [N+1], col [nnz])."""
    L = lib()
    conn = np.ascontiguousarray(mesh.conn, np.int32)
    rp = np.zeros(mesh.n_nodes + 1, np.int32)
    nnz = L.ox_ins_pattern(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(conn, _ip), _p(rp, _ip), None)
    col = np.zeros(nnz, np.int32)
    L.ox_ins_pattern(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(conn, _ip), _p(rp, _ip), _p(col, _ip))
    return rp, col


def ins_prescribe(vel_type, t, coords, node_vel=None):
    c = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    out = np.zeros_like(c)
    nv = None if node_vel is None else np.ascontiguousarray(node_vel, np.float64).reshape(-1, 3)
    r = lib().ox_ins_prescribe(int(vel_type), float(t), len(c), _p(c), None if nv is None else _p(nv), _p(out))
    if r != 0:
        raise ValueError("unknown velocity type or missing node velocities")
    return out


def ins_element(xel, vel, phi, phid, diff, dt, am, af, gm):
    """Element matrix and vector of the transient advection-diffusion stage."""
    xel = np.ascontiguousarray(xel, np.float64).reshape(-1, 3)
    npe = len(xel)
    vel = np.ascontiguousarray(vel, np.float64).reshape(npe, 3)
    phi = np.ascontiguousarray(phi, np.float64)
    phid = np.ascontiguousarray(phid, np.float64)
    diff = np.ascontiguousarray(diff, np.float64)
    emat = np.zeros((npe, npe))
    erhs = np.zeros(npe)
    lib().ox_ins_element(npe, _p(xel), _p(vel), _p(phi), _p(phid), _p(diff), dt, am, af, gm, _p(emat), _p(erhs))
    return emat, erhs


def gmres(row_ptr, col, val, b, tol=1e-5, max_iter=1000, restart=30):
    """Returns (x, converged, iterations, relative residual)."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    cl = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    it = np.zeros(1, np.int32)
    rr = np.zeros(1)
    ok = lib().ox_gmres(len(b), _p(rp, _ip), _p(cl, _ip), _p(v), _p(b), _p(x), tol, max_iter, restart,
                        _p(it, _ip), _p(rr))
    return x, bool(ok), int(it[0]), float(rr[0])


class INSOracle:
    """FP64 restatement of femIncompressibleSolver's advection-diffusion
    stage: assemble, Dirichlet rows, GMRES, generalized-alpha update."""

    def __init__(self, mesh, el_diff, dir_node, dir_val, dt, alpha_m, alpha_f, gamma, tol=1e-5, max_iter=1000,
                 restart=30):
        self.mesh = mesh
        self.k = _Keep(mesh, __import__("femorph_amd.mesh", fromlist=["BoundaryConditions"]).BoundaryConditions())
        self.rp, self.col = ins_pattern(mesh)
        N = mesh.n_nodes
        self.diff = np.ascontiguousarray(el_diff, np.float64).reshape(mesh.n_elems, 3)
        self.dir_node = np.ascontiguousarray(dir_node, np.int32)
        self.dir_val = np.ascontiguousarray(dir_val, np.float64).reshape(-1, 3)
        self.dt, self.am, self.af, self.gm = float(dt), float(alpha_m), float(alpha_f), float(gamma)
        self.tol, self.max_iter, self.restart = tol, max_iter, restart
        self.phi = np.zeros(N)
        self.phid = np.zeros(N)
        self.vel = np.zeros((N, 3))
        self.val = np.zeros(len(self.col))
        self.rhs = np.zeros(N)
        self.phid_new = None

    def prescribe(self, vel_type, t, node_vel=None):
        self.vel = ins_prescribe(vel_type, t, self.mesh.coords, node_vel)

    def assemble(self):
        r = lib().ox_ins_assemble(C.byref(self.k.m), _p(self.rp, _ip), _p(self.col, _ip), _p(self.vel),
                                  _p(self.phi), _p(self.phid), _p(self.diff), self.dt, self.am, self.af, self.gm,
                                  len(self.dir_node), _p(self.dir_node, _ip), _p(self.dir_val), _p(self.val),
                                  _p(self.rhs))
        if r != 0:
            raise RuntimeError("pattern mismatch")
        return self.val, self.rhs

    def advdiff_solve(self):
        self.assemble()
        x, ok, it, rr = gmres(self.rp, self.col, self.val, self.rhs, self.tol, self.max_iter, self.restart)
        self.phid_new = x
        return ok, it, rr

    def update(self):
        lib().ox_ins_update(len(self.phi), self.dt, self.gm, _p(self.phi), _p(self.phid), _p(self.phid_new))
