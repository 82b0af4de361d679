"""femx_ffd_morph_noalloc, the entry point that allocates nothing (the FFD
morph of device-resident points, its control-point lattice in a fixed
device array), against femx_ffd_morph through the C ABI, bit for bit.

Two calls in a row on one GPU share the array: the first with the largest
lattice (13 x 4 x 6 control points) and 64 displaced points, the second
with fewer points, a 3 x 2 x 3 lattice and 5 displaced points, so a stale
entry of the first call's lattice would show in the second's result.  The
test is the caller: it allocates the device buffers with the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX_A = ((0.1, -0.2, 0.05), (13, 4, 6), (0.9, 0.1, 0.0), (0.05, 1.1, 0.1), (0.0, -0.1, 0.8))
BOX_B = ((0.0, 0.0, 0.0), (3, 2, 3), (1.2, 0.0, 0.0), (0.0, 0.7, 0.0), (0.0, 0.0, 1.3))


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p

    def download(self, p, shape):
        out = np.empty(shape, np.float64)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(box, n_points, n_disp, seed):
    X0, n, S, T, U = box
    tot = n[0] * n[1] * n[2]
    rng = np.random.default_rng(seed)
    nodes = rng.integers(0, tot, n_disp)  # repeats allowed: the lattice adds them in list order
    vals = rng.uniform(-0.2, 0.2, (n_disp, 3))
    pts = rng.uniform(-0.3, 1.4, (n_points, 3))  # inside and outside the box
    return X0, n, S, T, U, nodes, vals, pts


def test_noalloc_equals_morph_on_two_calls(gpu):
    hip = _Hip()
    try:
        for box, n_points, n_disp, seed in ((BOX_A, 20000, 64, 1), (BOX_B, 3000, 5, 2)):
            case = _case(box, n_points, n_disp, seed)
            print(f"noalloc: lattice {box[1]}, {n_points} points, {n_disp} displaced", flush=True)
            pts = case[-1]
            dc = hip.upload(pts)
            dd = hip.alloc(pts.nbytes)
            assert hip.L.hipMemset(dd, 0xFF, C.c_size_t(pts.nbytes)) == 0  # NaN until written
            gpu.ffd_morph_noalloc(*case[:-1], n_points, dc.value, dd.value, None)
            assert hip.L.hipDeviceSynchronize() == 0
            got = hip.download(dd, pts.shape)
            ref = gpu.ffd_morph(*case)
            assert np.array_equal(got, ref), np.abs(got - ref).max()
    finally:
        hip.free()


def test_noalloc_rejects_sizes_beyond_its_limit(gpu):
    X0, n, S, T, U = BOX_B
    with pytest.raises(gpu.FemxError, match="at most 64"):
        gpu.ffd_morph_noalloc(X0, n, S, T, U, np.zeros(65, np.int32), np.zeros((65, 3)), 0, 0, 0)
    with pytest.raises(gpu.FemxError, match="2..13"):
        gpu.ffd_morph_noalloc(X0, (14, 2, 2), S, T, U, [0], [[0.1, 0.0, 0.0]], 0, 0, 0)
    with pytest.raises(gpu.FemxError, match="out of range"):
        gpu.ffd_morph_noalloc(X0, n, S, T, U, [18], [[0.1, 0.0, 0.0]], 0, 0, 0)
