"""Device memory cache of destroyed solver handles (femx_set_device_cache):
a destroyed VMS or SUPG handle returns its buffers to the cache, a new
handle of the same sizes is served from it, and its results are bit for bit
those of a handle on fresh allocations (the reused buffers hold the previous
solve's values, which set-up must overwrite)."""
import gc

import numpy as np
import pytest

from femorph_amd import meshgen

pytestmark = pytest.mark.gpu

CAP = 8 << 30


def _vms_run(gpu, m, bc, ic, steps):
    dm = gpu.DeviceMesh(m)
    g = gpu.VMSSolver(dm, bc, rho=1.0, mu=0.01, inv_eps=1.0, c1=4.0, c2=2.0, dt=0.1 * meshgen.min_tet_h(m),
                      precision=gpu.FEMX_MIXED)
    g.set_state(ic)
    g.step(steps)
    sol, sdus, t = g.get_state()
    del g, dm
    gc.collect()
    return sol, sdus, t


def _supg_run(gpu, m, bc, steps):
    dm = gpu.DeviceMesh(m)
    g = gpu.SUPGSolver(dm, bc, rho=1.0, mu=0.01, alpha=1.0, cI=4.0, dt=0.1 / 6, precision=gpu.FEMX_MIXED)
    g.set_state(np.zeros((m.n_nodes, 3)))
    g.step(steps)
    st = g.get_state()
    sol = np.concatenate([st["sol"].ravel(), st["pres"]])
    del g, dm
    gc.collect()
    return sol


def test_cached_buffers_give_the_same_results(gpu):
    try:
        gpu.set_device_cache(0)
        assert gpu.device_cache_bytes(0) == 0
        m = meshgen.tet_cube(6, 0.15, 3)
        bc = meshgen.lid_cavity_bc(m)
        ic = meshgen.random_velocity(m, 1e-2, 5)
        print("VMS 6^3x6 without the cache", flush=True)
        ref = _vms_run(gpu, m, bc, ic, 12)
        assert gpu.device_cache_bytes(0) == 0
        gpu.set_device_cache(CAP)
        print("VMS twice with the cache", flush=True)
        first = _vms_run(gpu, m, bc, ic, 12)
        held = gpu.device_cache_bytes(0)
        assert held > 0
        # a different state first, so that reused buffers hold other values
        other = _vms_run(gpu, m, bc, 0.5 * ic, 3)
        again = _vms_run(gpu, m, bc, ic, 12)
        assert gpu.device_cache_bytes(0) == held
        for r in (first, other, again):
            assert np.isfinite(r[0]).all()
        for a, b in zip(ref[:2], first[:2]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(ref[:2], again[:2]):
            np.testing.assert_array_equal(a, b)
        assert ref[2] == first[2] == again[2]

        h = meshgen.hex_cube(6, 0.2, 4)
        hbc = meshgen.lid_cavity_bc(h)
        print("SUPG 6^3 without and with the cache", flush=True)
        gpu.set_device_cache(0)
        sref = _supg_run(gpu, h, hbc, 10)
        gpu.set_device_cache(CAP)
        s1 = _supg_run(gpu, h, hbc, 10)
        s2 = _supg_run(gpu, h, hbc, 10)
        np.testing.assert_array_equal(sref, s1)
        np.testing.assert_array_equal(sref, s2)
        gpu.set_device_cache(0)
        assert gpu.device_cache_bytes(0) == 0
    finally:
        gpu.set_device_cache(CAP)
