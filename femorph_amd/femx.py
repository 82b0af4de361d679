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

# FEMX_MIXED: FP32 element math, FP64 residual accumulation, state and update.
# FEMX_MIXED_F32SUM (opt-in): as FEMX_MIXED with FP32 in-block node sums and FP32
# block partials; the sum over blocks, the state and the update stay FP64.
FEMX_FP64, FEMX_MIXED, FEMX_MIXED_F32SUM = 0, 1, 2
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


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("steps", C.c_int), ("mesh", MeshDesc), ("bc", BCDesc), ("vms", VMSParams),
                ("supg", SUPGParams)]


class RequestStats(C.Structure):
    _fields_ = [("id", C.c_longlong), ("kind", C.c_int), ("status", C.c_int), ("steps", C.c_int),
                ("steps_done", C.c_int), ("elements", C.c_longlong), ("nodes", C.c_longlong), ("bytes", C.c_longlong),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


class FFDBox(C.Structure):
    _fields_ = [("min_point", C.c_double * 3), ("tot_points", C.c_int * 3), ("axis_s", C.c_double * 3),
                ("axis_t", C.c_double * 3), ("axis_u", C.c_double * 3), ("n_disp", C.c_int),
                ("disp_node", _ip), ("disp_val", _dp)]


class INSParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("alpha_m", C.c_double), ("alpha_f", C.c_double), ("gamma", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("restart", C.c_int)]


class INSStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("rel_residual", C.c_double)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int), ("total_ms", C.c_double),
                ("bytes_per_launch", C.c_double)]


# (name, restype, argtypes) of every symbol of include/femx.h
SIGNATURES = [
    ("femx_last_error", C.c_char_p, []),
    ("femx_device_count", C.c_int, []),
    ("femx_set_default_precision", C.c_int, [C.c_int]),
    ("femx_default_precision", C.c_int, []),
    ("femx_set_default_linear_solve", C.c_int, [C.c_double, C.c_int, C.c_int]),
    ("femx_default_linear_solve", None, [_dp, _ip, _ip]),
    ("femx_set_device_cache", C.c_int, [C.c_size_t]),
    ("femx_device_cache_bytes", C.c_size_t, [C.c_int]),
    ("femx_mesh_upload", C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    ("femx_mesh_free", C.c_int, [C.c_void_p]),
    ("femx_mesh_block_stats", C.c_int, [C.POINTER(MeshDesc), C.c_int, C.POINTER(C.c_longlong)]),
    ("femx_mesh_block_stats_order", C.c_int, [C.POINTER(MeshDesc), C.c_int, C.c_int, C.POINTER(C.c_longlong)]),
    ("femx_mesh_upload_blocks", C.c_int, [C.POINTER(MeshDesc), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("femx_vms_create", C.c_int, [C.c_void_p, C.POINTER(BCDesc), C.POINTER(VMSParams), C.POINTER(C.c_void_p)]),
    ("femx_vms_destroy", C.c_int, [C.c_void_p]),
    ("femx_vms_set_state", C.c_int, [C.c_void_p, _dp, _dp]),
    ("femx_vms_step", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_vms_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("femx_vms_get_mass", C.c_int, [C.c_void_p, _dp]),
    ("femx_vms_get_cfl", C.c_int, [C.c_void_p, _dp, _dp]),
    ("femx_vms_set_diagnostics", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_vms_get_diagnostics", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("femx_vms_get_norms", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("femx_vms_sync", C.c_int, [C.c_void_p]),
    ("femx_vms_stream", C.c_void_p, [C.c_void_p]),
    ("femx_vms_union_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("femx_vms_union_destroy", C.c_int, [C.c_void_p]),
    ("femx_vms_union_add", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("femx_vms_union_step", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_vms_union_left", C.c_int, [C.c_void_p, C.c_void_p]),
    ("femx_vms_union_remove", C.c_int, [C.c_void_p, C.c_void_p]),
    ("femx_vms_union_sync", C.c_int, [C.c_void_p]),
    ("femx_vms_union_stream", C.c_void_p, [C.c_void_p]),
    ("femx_vms_batch_create", C.c_int, [C.c_int, C.POINTER(MeshDesc), C.POINTER(BCDesc), C.POINTER(VMSParams),
                                        C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("femx_vms_batch_destroy", C.c_int, [C.c_void_p]),
    ("femx_vms_batch_set_state", C.c_int, [C.c_void_p, C.POINTER(_dp)]),
    ("femx_vms_batch_step", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("femx_vms_batch_get_state", C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    ("femx_vms_batch_info", C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    ("femx_vms_batch_sync", C.c_int, [C.c_void_p]),
    ("femx_vms_batch_stream", C.c_void_p, [C.c_void_p]),
    ("femx_vms_batch_autotune", C.c_int, [C.c_int, C.POINTER(MeshDesc), C.POINTER(BCDesc), C.POINTER(VMSParams),
                                          C.c_int, _ip, _ip, _dp]),
    ("femx_ffd_control_points", C.c_int, [C.POINTER(FFDBox), _dp, _dp]),
    ("femx_ffd_morph", C.c_int, [C.POINTER(FFDBox), C.c_int, _dp, _dp]),
    ("femx_ffd_morph_noalloc", C.c_int, [C.POINTER(FFDBox), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("femx_supg_step_async", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_supg_finish", C.c_int, [C.c_void_p, _ip]),
    ("femx_supg_union_accepts", C.c_int, [C.c_void_p]),
    ("femx_supg_union_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("femx_supg_union_destroy", C.c_int, [C.c_void_p]),
    ("femx_supg_union_add", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("femx_supg_union_step", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_supg_union_left", C.c_int, [C.c_void_p, C.c_void_p]),
    ("femx_supg_union_remove", C.c_int, [C.c_void_p, C.c_void_p]),
    ("femx_supg_union_sync", C.c_int, [C.c_void_p]),
    ("femx_supg_union_stream", C.c_void_p, [C.c_void_p]),
    ("femx_server_create", C.c_int, [C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("femx_server_submit", C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_longlong)]),
    ("femx_server_poll", C.c_int, [C.c_void_p, C.POINTER(RequestStats), C.c_int]),
    ("femx_server_result", C.c_int, [C.c_void_p, C.c_longlong, _dp, C.c_longlong]),
    ("femx_server_drain", C.c_int, [C.c_void_p, C.c_double]),
    ("femx_server_info", C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    ("femx_server_destroy", C.c_int, [C.c_void_p]),
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
    ("femx_ins_create", C.c_int, [C.c_void_p, _dp, C.c_int, _ip, _dp, C.POINTER(INSParams), C.POINTER(C.c_void_p)]),
    ("femx_ins_destroy", C.c_int, [C.c_void_p]),
    ("femx_ins_set_state", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("femx_ins_prescribe_velocity", C.c_int, [C.c_void_p, C.c_int, C.c_double, _dp]),
    ("femx_ins_advdiff_solve", C.c_int, [C.c_void_p, C.POINTER(INSStats)]),
    ("femx_ins_update", C.c_int, [C.c_void_p]),
    ("femx_ins_get_state", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("femx_ins_get_system", C.c_int, [C.c_void_p, _ip, _ip, _dp, _dp, _dp, C.POINTER(C.c_longlong)]),
    ("femx_ins_sync", C.c_int, [C.c_void_p]),
    ("femx_ins_stream", C.c_void_p, [C.c_void_p]),
    ("femx_ins_kernel_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("femx_ins_kernel_times", C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int]),
    ("femx_hbm_probe", C.c_int, [C.c_int, C.c_longlong, C.c_int, C.c_int, _dp]),
    ("femx_valu_probe", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
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
        missing = []
        for name, res, args in SIGNATURES:
            try:
                f = getattr(L, name)
            except AttributeError:
                missing.append(name)
                continue
            f.restype = res
            f.argtypes = args
        # the in-tree library must export every entry point; an explicitly
        # chosen other build (FEMX_LIB, e.g. an older one being compared)
        # may lack newer ones, which then fail when called
        if missing and "FEMX_LIB" not in os.environ:
            raise RuntimeError(f"{LIB_PATH} lacks {missing}: rebuild it with `python -m femorph_amd.build`")
        _LIB = L
    return _LIB


def _check(status):
    if status != 0:
        raise FemxError(status, load().femx_last_error().decode())


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


BLOCKS_MORTON, BLOCKS_ELEMENT_ORDER = 0, 1


def block_stats(mesh: Mesh, block_elems: int, order: int = BLOCKS_MORTON) -> dict:
    """Host-only statistics of the element-block partition (no GPU needed);
    order BLOCKS_MORTON (femx_mesh_upload's) or BLOCKS_ELEMENT_ORDER."""
    coords = np.ascontiguousarray(mesh.coords, np.float64)
    conn = np.ascontiguousarray(mesh.conn, np.int32)
    d = MeshDesc(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(coords), _p(conn, _ip))
    st = (C.c_longlong * 8)()
    _check(load().femx_mesh_block_stats_order(C.byref(d), block_elems, order, st))
    keys = ["blocks", "elements", "local_nodes", "max_local", "shared_nodes", "partials", "complete", "max_elems"]
    return dict(zip(keys, [int(x) for x in st]))


def device_count() -> int:
    return load().femx_device_count()


def set_device_cache(max_bytes_per_device: int) -> None:
    """Cap of the per-device cache that destroyed solver handles return
    their device buffers and streams to (femx_set_device_cache); 0 frees
    what it holds and turns it off."""
    _check(load().femx_set_device_cache(int(max_bytes_per_device)))


def device_cache_bytes(device: int = 0) -> int:
    """Bytes of idle device memory the cache holds on a device."""
    return int(load().femx_device_cache_bytes(int(device)))


HBM_PROBE_KINDS = {"copy": 0, "read": 1, "write": 2, "mix_7r5w": 3,
                   "copy_nt": 4, "read_nt": 5, "write_nt": 6, "mix_7r5w_nt": 7}


def hbm_probe(kind: str, nbytes: int = 1 << 30, grid: int = 4096, reps: int = 20) -> float:
    """TB/s of a float4 streaming kernel (femx_hbm_probe): copy, read,
    write, or a 7-read / 5-write mix, over buffers of nbytes; the "_nt"
    kinds use non-temporal loads and stores."""
    out = C.c_double(0.0)
    _check(load().femx_hbm_probe(HBM_PROBE_KINDS[kind], int(nbytes), int(grid), int(reps), C.byref(out)))
    return out.value


def valu_probe(kind: str, grid: int = 8192, iters: int = 4096, reps: int = 10) -> float:
    """TFLOP/s of eight independent FMA chains per thread (femx_valu_probe):
    kind "f32" (v_fma_f32, no packed FP32) or "f64" (v_fma_f64)."""
    out = C.c_double(0.0)
    _check(load().femx_valu_probe({"f32": 0, "f64": 1}[kind], int(grid), int(iters), int(reps), C.byref(out)))
    return out.value


class DeviceMesh:
    """A mesh uploaded to the GPU (femx_mesh_upload_blocks): element blocks
    of block_elems (0: the default) in Morton order or in the mesh's element
    order (BLOCKS_ELEMENT_ORDER)."""

    def __init__(self, mesh: Mesh, block_elems: int = 0, order: int = BLOCKS_MORTON):
        L = load()
        self.mesh = mesh
        self._coords = np.ascontiguousarray(mesh.coords, np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, np.int32)
        d = MeshDesc(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(self._coords), _p(self._conn, _ip))
        h = C.c_void_p()
        _check(L.femx_mesh_upload_blocks(C.byref(d), int(block_elems), int(order), C.byref(h)))
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


class _BatchDescs:
    """C descriptors of a list of (mesh, bc, params dict) kept alive together."""

    def __init__(self, items):
        n = len(items)
        self.keep = []
        self.meshes = (MeshDesc * n)()
        self.bcs = (BCDesc * n)()
        self.params = (VMSParams * n)()
        for i, (mesh, bc, prm) in enumerate(items):
            if mesh.npe != 4:
                raise ValueError("batches hold tet4 meshes")
            c = np.ascontiguousarray(mesh.coords, np.float64)
            e = np.ascontiguousarray(mesh.conn, np.int32)
            k = _BCKeep(bc)
            self.keep += [c, e, k]
            self.meshes[i] = MeshDesc(mesh.n_nodes, mesh.n_elems, 4, _p(c), _p(e, _ip))
            self.bcs[i] = k.desc
            dt = prm["dt"]
            self.params[i] = VMSParams(prm["rho"], prm["mu"], prm["inv_eps"], prm["c1"], prm["c2"], dt,
                                       prm.get("dual_step", dt), prm.get("tot_dual_steps", 0),
                                       prm.get("precision", FEMX_FP64))


def vms_batch_autotune(items, steps=6):
    """Time block sizes {256, 224, 192} x xcd order {1, 0} on `items`;
    returns (block_elems, xcd_order, ms_per_step[6])."""
    d = _BatchDescs(items)
    b, x, ms = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(6)
    _check(load().femx_vms_batch_autotune(len(items), d.meshes, d.bcs, d.params, steps, _p(b, _ip), _p(x, _ip),
                                          _p(ms)))
    return int(b[0]), int(x[0]), ms


class VMSUnion:
    """Continuous batching of VMS handles (femx_vms_union_*): members of any
    sizes advanced together, one launch per phase per step, each by its own
    number of steps; bit for bit the handles stepped on their own."""

    def __init__(self):
        h = C.c_void_p()
        _check(load().femx_vms_union_create(C.byref(h)))
        self._h = h
        self._members = {}

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_vms_union_destroy(self._h)
            self._h = None

    def add(self, solver, steps):
        _check(load().femx_vms_union_add(self._h, solver._h, int(steps)))
        self._members[id(solver)] = solver

    def step(self, k):
        _check(load().femx_vms_union_step(self._h, int(k)))

    def left(self, solver):
        return load().femx_vms_union_left(self._h, solver._h)

    def remove(self, solver):
        _check(load().femx_vms_union_remove(self._h, solver._h))
        self._members.pop(id(solver), None)

    def sync(self):
        _check(load().femx_vms_union_sync(self._h))


class SUPGUnion:
    """Continuous batching of SUPG handles (femx_supg_union_*): FEMX_MIXED
    hexa8 members advanced together, one launch per phase per step, each by
    its own number of steps; bit for bit the handles stepped on their own.
    After remove(), the member's finish() checks its steps for the exit
    condition."""

    def __init__(self):
        h = C.c_void_p()
        _check(load().femx_supg_union_create(C.byref(h)))
        self._h = h
        self._members = {}

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_supg_union_destroy(self._h)
            self._h = None

    def add(self, solver, steps):
        _check(load().femx_supg_union_add(self._h, solver._h, int(steps)))
        self._members[id(solver)] = solver

    def step(self, k):
        _check(load().femx_supg_union_step(self._h, int(k)))

    def left(self, solver):
        return load().femx_supg_union_left(self._h, solver._h)

    def remove(self, solver):
        _check(load().femx_supg_union_remove(self._h, solver._h))
        self._members.pop(id(solver), None)

    def sync(self):
        _check(load().femx_supg_union_sync(self._h))


class VMSBatch:
    """Many VMS tet4 solves advanced together (femx_vms_batch_*).

    items: list of (Mesh, BoundaryConditions, dict(rho, mu, inv_eps, c1, c2,
    dt[, dual_step, tot_dual_steps, precision]))."""

    def __init__(self, items, block_elems=0, xcd_order=1):
        self.sizes = [m.n_nodes for m, _, _ in items]
        d = _BatchDescs(items)
        h = C.c_void_p()
        _check(load().femx_vms_batch_create(len(items), d.meshes, d.bcs, d.params, block_elems, xcd_order,
                                            C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_vms_batch_destroy(self._h)
            self._h = None

    def set_state(self, sols=None):
        if sols is None:
            _check(load().femx_vms_batch_set_state(self._h, None))
            return
        arrs = [None if s is None else np.ascontiguousarray(s, np.float64).reshape(n, 4)
                for s, n in zip(sols, self.sizes)]
        ptrs = (_dp * len(arrs))(*[_p(a) if a is not None else None for a in arrs])
        self._keep = arrs
        _check(load().femx_vms_batch_set_state(self._h, ptrs))

    def step(self, n=1, graph=True):
        _check(load().femx_vms_batch_step(self._h, n, int(graph)))

    def get_state(self, i):
        sol, t = np.zeros((self.sizes[i], 4)), np.zeros(1)
        _check(load().femx_vms_batch_get_state(self._h, i, _p(sol), _p(t)))
        return sol, float(t[0])

    def info(self):
        a = (C.c_longlong * 8)()
        _check(load().femx_vms_batch_info(self._h, a))
        return dict(zip(("meshes", "nodes", "elements", "blocks", "shared_nodes", "block_elems", "xcd_order",
                         "max_local"), list(a)))

    def sync(self):
        _check(load().femx_vms_batch_sync(self._h))

    def stream(self):
        return load().femx_vms_batch_stream(self._h)


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
        """0/False off, 1/True norms and element fields, 2 norms only."""
        _check(load().femx_vms_set_diagnostics(self._h, int(enable)))

    def get_norms(self, nsteps):
        """Per-iteration norms of the last nsteps (<= 64) steps run with
        diagnostics, oldest first: [nsteps, tot_dual_steps + 1, 3]."""
        out = np.zeros((nsteps, self.tot_dual_steps + 1, 3))
        _check(load().femx_vms_get_norms(self._h, int(nsteps), _p(out)))
        return out

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

    def step_async(self, n=1):
        """Enqueues n steps without waiting (femx_supg_step_async)."""
        _check(load().femx_supg_step_async(self._h, n))

    def finish(self):
        """Waits for the enqueued steps (also a union member's, after it
        left) and returns the number done (femx_supg_finish)."""
        done = C.c_int(0)
        st = load().femx_supg_finish(self._h, C.byref(done))
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


class INSSolver:
    """Implicit transient advection-diffusion stage of femIncompressibleSolver
    (assembly, Dirichlet rows, GMRES, generalized-alpha update) on the GPU."""

    def __init__(self, dmesh: DeviceMesh, dt, alpha_m, alpha_f, gamma, el_diff=None, dir_node=(), dir_val=None,
                 tol=1.0e-5, max_iter=1000, restart=30):
        L = load()
        self.dmesh = dmesh
        self.N, self.E = dmesh.mesh.n_nodes, dmesh.mesh.n_elems
        dn = np.ascontiguousarray(dir_node, np.int32)
        dv = np.ascontiguousarray(np.zeros((len(dn), 3)) if dir_val is None else dir_val, np.float64).reshape(-1, 3)
        ed = None if el_diff is None else np.ascontiguousarray(el_diff, np.float64).reshape(self.E, 3)
        p = INSParams(dt, alpha_m, alpha_f, gamma, tol, max_iter, restart)
        h = C.c_void_p()
        _check(L.femx_ins_create(dmesh._h, None if ed is None else _p(ed), len(dn), _p(dn, _ip), _p(dv),
                                 C.byref(p), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_ins_destroy(self._h)
            self._h = None

    def set_state(self, phi=None, phi_dot=None, vel=None):
        a = [None if x is None else np.ascontiguousarray(x, np.float64).reshape(self.N, k)
             for x, k in ((phi, 1), (phi_dot, 1), (vel, 3))]
        _check(load().femx_ins_set_state(self._h, *[None if x is None else _p(x) for x in a]))

    def prescribe_velocity(self, vel_type, t, node_vel=None):
        nv = None if node_vel is None else np.ascontiguousarray(node_vel, np.float64).reshape(self.N, 3)
        _check(load().femx_ins_prescribe_velocity(self._h, int(vel_type), float(t), None if nv is None else _p(nv)))

    def advdiff_solve(self):
        st = INSStats()
        _check(load().femx_ins_advdiff_solve(self._h, C.byref(st)))
        return dict(iterations=st.iterations, converged=bool(st.converged), rel_residual=st.rel_residual)

    def update(self):
        _check(load().femx_ins_update(self._h))

    def get_state(self):
        phi, phid, vel = np.zeros(self.N), np.zeros(self.N), np.zeros((self.N, 3))
        _check(load().femx_ins_get_state(self._h, _p(phi), _p(phid), _p(vel)))
        return phi, phid, vel

    def get_system(self):
        """(row_ptr, col, val, rhs, phid_new) of the last solve."""
        nnz = C.c_longlong(0)
        L = load()
        _check(L.femx_ins_get_system(self._h, None, None, None, None, None, C.byref(nnz)))
        rp, col = np.zeros(self.N + 1, np.int32), np.zeros(nnz.value, np.int32)
        val, rhs, x = np.zeros(nnz.value), np.zeros(self.N), np.zeros(self.N)
        _check(L.femx_ins_get_system(self._h, _p(rp, _ip), _p(col, _ip), _p(val), _p(rhs), _p(x), None))
        return rp, col, val, rhs, x

    def sync(self):
        _check(load().femx_ins_sync(self._h))

    def stream(self):
        return load().femx_ins_stream(self._h)

    def kernel_timing(self, enable=True):
        _check(load().femx_ins_kernel_timing(self._h, int(enable)))

    def kernel_times(self):
        return _kernel_times(load().femx_ins_kernel_times, self._h)


class Server:
    """Solve server (femx_server_*): admission by device footprint, resident
    state, every resident request advanced on its own stream."""

    STATS = ("id", "kind", "status", "steps", "steps_done", "elements", "nodes", "bytes",
             "t_submit", "t_admit", "t_start", "t_done")

    def __init__(self, budget_bytes=64 << 30, max_active=32, prep_threads=4, chunk_steps=25, gpu_threads=4):
        h = C.c_void_p()
        _check(load().femx_server_create(int(budget_bytes), max_active, prep_threads, gpu_threads, chunk_steps,
                                         C.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.femx_server_destroy(self._h)
            self._h = None

    __del__ = close

    def _submit(self, kind, mesh, bc, steps, vms=None, supg=None):
        c = np.ascontiguousarray(mesh.coords, np.float64)
        e = np.ascontiguousarray(mesh.conn, np.int32)
        k = _BCKeep(bc)
        r = Request(kind, steps, MeshDesc(mesh.n_nodes, mesh.n_elems, mesh.npe, _p(c), _p(e, _ip)), k.desc,
                    vms if vms is not None else VMSParams(), supg if supg is not None else SUPGParams())
        rid = C.c_longlong()
        _check(load().femx_server_submit(self._h, C.byref(r), C.byref(rid)))  # arrays are copied here
        self._n[rid.value] = (mesh.n_nodes, 4 if kind == 0 else 3)
        return rid.value

    def submit_vms(self, mesh, bc, steps, rho, mu, inv_eps, c1, c2, dt, dual_step=None, tot_dual_steps=0,
                   precision=FEMX_MIXED):
        p = VMSParams(rho, mu, inv_eps, c1, c2, dt, dt if dual_step is None else dual_step, tot_dual_steps, precision)
        return self._submit(0, mesh, bc, steps, vms=p)

    def submit_supg(self, mesh, bc, steps, rho, mu, alpha, cI, dt, exopts=(0, 1, 1, 1, 1, 1), fixed_dofs="NONE",
                    exit_condition=1.0e5, precision=FEMX_MIXED):
        e = [int(bool(x)) for x in exopts]
        if mesh.npe == 4:
            e[1] = 0  # B-hat is hexa8-only in the reference
        p = SUPGParams(rho, mu, alpha, cI, dt, e[0], e[1], e[2], e[3], e[4], e[5], FIXED_DOFS[fixed_dofs],
                       exit_condition, precision)
        return self._submit(1, mesh, bc, steps, supg=p)

    def poll(self, max_entries=4096):
        buf = (RequestStats * max_entries)()
        n = load().femx_server_poll(self._h, buf, max_entries)
        return [{f: getattr(buf[i], f) for f in self.STATS} for i in range(n)]

    def result(self, rid):
        """Final state of a finished request: VMS [N, 4], SUPG [N, 3]."""
        nn, dofs = self._n.pop(rid)
        out = np.zeros((nn, dofs))
        _check(load().femx_server_result(self._h, rid, _p(out), nn * dofs))
        return out

    def drain(self, timeout_s=600.0):
        _check(load().femx_server_drain(self._h, timeout_s))

    def info(self):
        a = (C.c_longlong * 7)()
        _check(load().femx_server_info(self._h, a))
        return dict(zip(("queued", "resident", "reserved", "peak_reserved", "budget", "max_active", "gpu_threads"),
                        list(a)))


def _ffd_box(min_point, n, S, T, U, disp_node, disp_val):
    dn = np.ascontiguousarray(disp_node, np.int32)
    dv = np.ascontiguousarray(disp_val, np.float64).reshape(-1, 3)
    b = FFDBox((C.c_double * 3)(*min_point), (C.c_int * 3)(*n), (C.c_double * 3)(*S), (C.c_double * 3)(*T),
               (C.c_double * 3)(*U), len(dn), _p(dn, _ip), _p(dv))
    return b, (dn, dv)


def ffd_morph(min_point, n, S, T, U, disp_node, disp_val, coords):
    """Free-form deformation of points by one Bernstein box on the GPU
    (femPointGrid::morphPoints); returns [n, 3] displacements (box
    coordinates for points outside the box, as the reference stores)."""
    b, keep = _ffd_box(min_point, n, S, T, U, disp_node, disp_val)
    c = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    out = np.zeros_like(c)
    _check(load().femx_ffd_morph(C.byref(b), len(c), _p(c), _p(out)))
    return out


def ffd_morph_noalloc(min_point, n, S, T, U, disp_node, disp_val, n_points, coords_dev, disp_dev, stream=None):
    """femx_ffd_morph_noalloc: the morph of n_points device-resident points
    (coords_dev, disp_dev: device addresses of [n][3] doubles), enqueued on
    `stream` (a hipStream_t address, None = the null stream) without any
    device allocation; the caller synchronises the stream."""
    b, keep = _ffd_box(min_point, n, S, T, U, disp_node, disp_val)
    _check(load().femx_ffd_morph_noalloc(C.byref(b), int(n_points), C.c_void_p(coords_dev), C.c_void_p(disp_dev),
                                         C.c_void_p(stream)))


def ffd_control_points(min_point, n, S, T, U, disp_node, disp_val):
    """Control-point coordinates and displacements of one box (host only)."""
    b, keep = _ffd_box(min_point, n, S, T, U, disp_node, disp_val)
    tot = int(n[0]) * int(n[1]) * int(n[2])
    c, d = np.zeros((tot, 3)), np.zeros((tot, 3))
    _check(load().femx_ffd_control_points(C.byref(b), _p(c), _p(d)))
    return c, d
