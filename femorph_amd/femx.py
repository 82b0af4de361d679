"""ctypes binding of the femx C ABI (include/femx.h) -> femorph_amd/libfemx.so.

Python mirror of the reference's solver interface for tests and benchmarks.
There is no CPU fallback: if the HIP library is not built, or no GPU is
visible, the constructors raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .mesh import BoundaryConditions, Mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
# FEMX_LIB selects another build of the same library (A/B comparisons).
LIB_PATH = os.environ.get("FEMX_LIB", os.path.join(_HERE, "libfemx.so"))

FEMX_FP64, FEMX_MIXED = 0, 1
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_BC", 4: "ERR_NOMEM", 5: "ERR_UNSUPPORTED",
          6: "EXIT_CONDITION", 7: "ERR_NO_DEVICE"}
EXIT_CONDITION = 6

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MeshDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_elems", C.c_int), ("npe", C.c_int), ("coords", _dp), ("conn", _ip)]


class BCDesc(C.Structure):
    _fields_ = [("n_vel", C.c_int), ("vel_node", _ip), ("vel_dir", _ip), ("vel_row_off", _ip), ("vel_rows", _dp),
                ("n_dir", C.c_int), ("dir_node", _ip), ("dir_val", _dp),
                ("n_pres", C.c_int), ("pres_node", _ip), ("pres_val", _dp)]


class VMSParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("inv_eps", C.c_double), ("c1", C.c_double),
                ("c2", C.c_double), ("dt", C.c_double), ("dual_step", C.c_double), ("tot_dual_steps", C.c_int),
                ("precision", C.c_int)]


class SUPGParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("alpha", C.c_double), ("cI", C.c_double),
                ("dt", C.c_double), ("use_euler", C.c_int), ("use_bhat", C.c_int), ("k1", C.c_int),
                ("k2", C.c_int), ("k3", C.c_int), ("k4", C.c_int), ("fixed_dofs", C.c_int),
                ("exit_condition", C.c_double), ("precision", C.c_int)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int), ("total_ms", C.c_double),
                ("bytes_per_launch", C.c_double)]


# (name, restype, argtypes) of every symbol of include/femx.h
SIGNATURES = [
    ("femx_last_error", C.c_char_p, []),
    ("femx_device_count", C.c_int, []),
    ("femx_mesh_upload", C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    ("femx_mesh_free", C.c_int, [C.c_void_p]),
    ("femx_mesh_block_stats", C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(C.c_longlong)]),
    ("femx_vms_create", C.c_int, [C.c_void_p, C.POINTER(BCDesc), C.POINTER(VMSParams), C.POINTER(C.c_void_p)]),
    ("femx_vms_destroy", C.c_int, [C.c_void_p]),
    ("femx_vms_set_state", C.c_int, [C.c_void_p, _dp, _dp]),
    ("femx_vms_step", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_vms_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("femx_vms_get_mass", C.c_int, [C.c_void_p, _dp]),
    ("femx_vms_get_cfl", C.c_int, [C.c_void_p, _dp, _dp]),
    ("femx_vms_set_diagnostics", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_vms_get_diagnostics", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("femx_vms_sync", C.c_int, [C.c_void_p]),
    ("femx_vms_stream", C.c_void_p, [C.c_void_p]),
    ("femx_supg_create", C.c_int, [C.c_void_p, C.POINTER(BCDesc), C.POINTER(SUPGParams), C.POINTER(C.c_void_p)]),
    ("femx_supg_destroy", C.c_int, [C.c_void_p]),
    ("femx_supg_set_state", C.c_int, [C.c_void_p, _dp]),
    ("femx_supg_step", C.c_int, [C.c_void_p, C.c_int, _ip]),
    ("femx_supg_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("femx_supg_get_mass", C.c_int, [C.c_void_p, _dp]),
    ("femx_supg_get_log", C.c_int, [C.c_void_p, _dp]),
    ("femx_supg_set_diagnostics", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_supg_sync", C.c_int, [C.c_void_p]),
    ("femx_supg_stream", C.c_void_p, [C.c_void_p]),
    ("femx_vms_kernel_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_vms_kernel_times", C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int]),
    ("femx_supg_kernel_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_supg_kernel_times", C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int]),
]

_LIB = None


class FemxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"femx {STATUS.get(status, status)}: {msg}")
        self.status = status


def load() -> C.CDLL:
    """Load libfemx.so (raises if it was not built: there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m femorph_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(status):
    if status != 0:
        raise FemxError(status, load().femx_last_error().decode())


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def block_stats(mesh: Mesh, block_elems: int) -> dict:
    """Host-only statistics of the element-block partition (no GPU needed)."""
    coords = np.ascontiguousarray(mesh.coords, np.float64)
    conn = np.ascontiguousarray(mesh.conn, np.int32)
    d = MeshDesc(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(coords), _p(conn, _ip))
    st = (C.c_longlong * 8)()
    _check(load().femx_mesh_block_stats(C.byref(d), block_elems, st))
    keys = ["blocks", "elements", "local_nodes", "max_local", "shared_nodes", "partials", "complete", "max_elems"]
    return dict(zip(keys, [int(x) for x in st]))


def device_count() -> int:
    return load().femx_device_count()


class DeviceMesh:
    """A mesh uploaded to the GPU (femx_mesh_upload)."""

    def __init__(self, mesh: Mesh):
        L = load()
        self.mesh = mesh
        self._coords = np.ascontiguousarray(mesh.coords, np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, np.int32)
        d = MeshDesc(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(self._coords), _p(self._conn, _ip))
        h = C.c_void_p()
        _check(L.femx_mesh_upload(C.byref(d), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_mesh_free(self._h)
            self._h = None


class _BCKeep:
    def __init__(self, bc: BoundaryConditions):
        self.a = [np.ascontiguousarray(x) for x in (bc.vel_node, bc.vel_dir, bc.vel_row_off, bc.vel_rows,
                                                     bc.dir_node, bc.dir_val, bc.pres_node, bc.pres_val)]
        a = self.a
        self.desc = BCDesc(bc.n_vel, _p(a[0], _ip), _p(a[1], _ip), _p(a[2], _ip), _p(a[3]),
                           len(bc.dir_node), _p(a[4], _ip), _p(a[5]), len(bc.pres_node), _p(a[6], _ip), _p(a[7]))


def _kernel_times(fn, h):
    buf = (KernelTime * 32)()
    n = fn(h, buf, 32)
    return {buf[i].name.decode(): dict(launches=buf[i].launches, total_ms=buf[i].total_ms,
                                       bytes_per_launch=buf[i].bytes_per_launch) for i in range(min(n, 32))}


class VMSSolver:
    """Device-resident explicit VMS solver (femVMSExplicitFluidSolver)."""

    def __init__(self, dmesh: DeviceMesh, bc: BoundaryConditions, rho, mu, inv_eps, c1, c2, dt,
                 dual_step=None, tot_dual_steps=0, precision=FEMX_FP64):
        L = load()
        self.dmesh = dmesh
        self.N, self.E = dmesh.mesh.n_nodes, dmesh.mesh.n_elems
        self.tot_dual_steps = tot_dual_steps
        k = _BCKeep(bc)
        p = VMSParams(rho, mu, inv_eps, c1, c2, dt, dt if dual_step is None else dual_step, tot_dual_steps, precision)
        h = C.c_void_p()
        _check(L.femx_vms_create(dmesh._h, C.byref(k.desc), C.byref(p), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_vms_destroy(self._h)
            self._h = None

    def set_state(self, sol, sdus=None):
        sol = np.ascontiguousarray(sol, np.float64).reshape(self.N, 4)
        sd = None
        if sdus is not None:
            sd = np.ascontiguousarray(sdus, np.float64).reshape(12 * self.E)
        _check(load().femx_vms_set_state(self._h, _p(sol), _p(sd) if sd is not None else None))

    def step(self, n=1):
        _check(load().femx_vms_step(self._h, n))

    def get_state(self):
        sol = np.zeros((self.N, 4))
        sdus = np.zeros((12, self.E))
        t = np.zeros(1)
        _check(load().femx_vms_get_state(self._h, _p(sol), _p(sdus), _p(t)))
        return sol, sdus, float(t[0])

    def get_mass(self):
        m = np.zeros(self.N)
        _check(load().femx_vms_get_mass(self._h, _p(m)))
        return m

    def get_cfl(self):
        a, b = np.zeros(1), np.zeros(1)
        _check(load().femx_vms_get_cfl(self._h, _p(a), _p(b)))
        return float(a[0]), float(b[0])

    def set_diagnostics(self, enable=True):
        _check(load().femx_vms_set_diagnostics(self._h, int(enable)))

    def get_diagnostics(self):
        """Element diagnostics of the last step: dict with iter_norms
        [(tot_dual_steps+1), 3] (max |res_u|, max |res_p|, CFL per dual
        iteration) and res_v [E,3], res_p [E], subgrid_u [E,3], subgrid_p [E]
        of the last dual iteration."""
        E = self.E
        it = np.zeros((self.tot_dual_steps + 1, 3))
        ru, rp, su, sp = np.zeros((E, 3)), np.zeros(E), np.zeros((E, 3)), np.zeros(E)
        _check(load().femx_vms_get_diagnostics(self._h, _p(it), _p(ru), _p(rp), _p(su), _p(sp)))
        return dict(iter_norms=it, res_v=ru, res_p=rp, subgrid_u=su, subgrid_p=sp)

    def sync(self):
        _check(load().femx_vms_sync(self._h))

    def stream(self):
        return load().femx_vms_stream(self._h)

    def kernel_timing(self, enable=True):
        _check(load().femx_vms_kernel_timing(self._h, int(enable)))

    def kernel_times(self):
        return _kernel_times(load().femx_vms_kernel_times, self._h)


FIXED_DOFS = {"NONE": 0, "X": 1, "Y": 2, "Z": 3, "XY": 4, "YZ": 5, "XZ": 6}


class SUPGSolver:
    """Device-resident explicit SUPG solver (femSUPGExplicitFluidSolver)."""

    def __init__(self, dmesh: DeviceMesh, bc: BoundaryConditions, rho, mu, alpha, cI, dt,
                 exopts=(0, 1, 1, 1, 1, 1), fixed_dofs="NONE", exit_condition=1.0e5, precision=FEMX_FP64):
        L = load()
        self.dmesh = dmesh
        self.N, self.E = dmesh.mesh.n_nodes, dmesh.mesh.n_elems
        k = _BCKeep(bc)
        e = [int(bool(x)) for x in exopts]
        p = SUPGParams(rho, mu, alpha, cI, dt, e[0], e[1], e[2], e[3], e[4], e[5], FIXED_DOFS[fixed_dofs],
                       exit_condition, precision)
        h = C.c_void_p()
        _check(L.femx_supg_create(dmesh._h, C.byref(k.desc), C.byref(p), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_supg_destroy(self._h)
            self._h = None

    def set_state(self, sol):
        sol = np.ascontiguousarray(np.asarray(sol, np.float64)[:, :3]).reshape(self.N, 3)
        _check(load().femx_supg_set_state(self._h, _p(sol)))

    def step(self, n=1):
        """Runs n steps; returns the number done (fewer if the exit condition fired)."""
        done = C.c_int(0)
        st = load().femx_supg_step(self._h, n, C.byref(done))
        if st == EXIT_CONDITION:
            return done.value
        _check(st)
        return done.value

    def get_state(self):
        sol = np.zeros((self.N, 3))
        pres, lam, reg = np.zeros(self.E), np.zeros(self.E), np.zeros(self.E)
        t = np.zeros(1)
        _check(load().femx_supg_get_state(self._h, _p(sol), _p(pres), _p(lam), _p(reg), _p(t)))
        return dict(sol=sol, pres=pres, lam=lam, regime=reg, time=float(t[0]))

    def get_mass(self):
        m = np.zeros(3 * self.N)
        _check(load().femx_supg_get_mass(self._h, _p(m)))
        return m

    def set_diagnostics(self, enable=True):
        _check(load().femx_supg_set_diagnostics(self._h, int(enable)))

    def get_log(self):
        v = np.zeros(19)
        _check(load().femx_supg_get_log(self._h, _p(v)))
        return v

    def sync(self):
        _check(load().femx_supg_sync(self._h))

    def stream(self):
        return load().femx_supg_stream(self._h)

    def kernel_timing(self, enable=True):
        _check(load().femx_supg_kernel_timing(self._h, int(enable)))

    def kernel_times(self):
        return _kernel_times(load().femx_supg_kernel_times, self._h)
