"""Pins of the oracle's integration rules, shape functions and kinematics
(src/femIntegrationRule.cpp, src/femHexa8.cpp:12-42, src/femTetra4.cpp:295-327,
src/femElement.cpp:104-354) against the compiled reference rule
(oracle/_ref, from the unmodified femIntegrationRule.cpp) and closed forms."""
import itertools

import numpy as np
import pytest

from oracle import oracle


def test_rules_match_compiled_reference():
    L = oracle.ref_rules_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent here)")
    import ctypes as C
    for npe in (4, 8):
        c = np.zeros(32)
        w = np.zeros(8)
        n = L.ref_rule(1, npe, 2, c.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)))
        oc, ow = oracle.rule_second_order(npe)
        assert n == len(ow)
        cols = 4 if npe == 4 else 3
        assert np.array_equal(c.reshape(8, 4)[:n, :cols], oc[:, :cols])
        assert np.array_equal(w[:n], ow)


def test_rule_closed_forms():
    c4, w4 = oracle.rule_second_order(4)
    assert abs(w4.sum() - 1.0 / 6.0) < 1e-16
    # literal 8-digit constants of the reference (they sum to 1 only to 1e-8)
    assert set(np.unique(c4)) == {0.58541020, 0.13819660}
    # degree-2 exactness on the reference tet, to the truncation of the constants
    for p in [(0, 0, 0), (1, 0, 0), (2, 0, 0), (1, 1, 0), (0, 1, 1)]:
        x, y, z = c4[:, 0], c4[:, 1], c4[:, 2]
        q = np.sum(w4 * x ** p[0] * y ** p[1] * z ** p[2])
        from math import factorial
        exact = factorial(p[0]) * factorial(p[1]) * factorial(p[2]) / factorial(sum(p) + 3)
        assert abs(q - exact) < 1e-8
    c8, w8 = oracle.rule_second_order(8)
    assert abs(w8.sum() - 8.0) < 1e-15
    # 2x2x2 Gauss is exact up to degree 3 per direction
    for p in itertools.product(range(4), repeat=3):
        q = np.sum(w8 * c8[:, 0] ** p[0] * c8[:, 1] ** p[1] * c8[:, 2] ** p[2])
        exact = np.prod([0.0 if k % 2 else 2.0 / (k + 1) for k in p])
        assert abs(q - exact) < 1e-14
    # order of the reference: the zeta = + layer first, counter-clockwise
    s = np.sign(c8[:, :3])
    assert s.tolist() == [[-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                          [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]]


@pytest.mark.parametrize("npe", [4, 8])
def test_partition_of_unity_and_gradients(npe):
    rng = np.random.default_rng(0)
    c, _ = oracle.rule_second_order(npe)
    if npe == 8:
        ref = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    else:  # femTetra4 node order with N3 = 1 - sum(xi)
        ref = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]], float)
    xel = ref * [1.3, 0.8, 1.1] + rng.uniform(-0.1, 0.1, ref.shape)
    f = 2 * xel[:, 0] - xel[:, 1] + xel[:, 2]  # femTestSolver field (src/femSolver.cpp:689-718)
    for g in range(len(c)):
        N = oracle.shape(npe, c[g, :3])
        assert abs(N.sum() - 1.0) < 1e-15
        det, DN = oracle.global_deriv(npe, xel, c[g, :3])
        assert np.abs(DN.sum(axis=0)).max() < 1e-13
        np.testing.assert_allclose(DN.T @ f, [2.0, -1.0, 1.0], rtol=0, atol=1e-13)
        # metric G = J^-T J^-1 in the reference's index order (G[a][b] = sum_c iJ[c][a] iJ[c][b])
        G = oracle.geom_matrix(npe, xel, c[g, :3])
        h = 1e-6
        J = np.zeros((3, 3))
        for a in range(3):
            xp = c[g, :3].copy(); xp[a] += h
            xm = c[g, :3].copy(); xm[a] -= h
            J[a] = (oracle.shape(npe, xp) @ xel - oracle.shape(npe, xm) @ xel) / (2 * h)
        iJ = np.linalg.inv(J)
        np.testing.assert_allclose(G, iJ.T @ iJ, rtol=1e-7)
        assert abs(det - np.linalg.det(J)) < 1e-7 * abs(det)


def test_interp_table():
    rows = np.array([[0.0, 1.0, 2.0, 0.0], [2.0, 3.0, 6.0, 1.0], [4.0, 4.0, 4.0, 3.0]])
    np.testing.assert_allclose(oracle.interp_table(0.5, rows), [1.0, 2.0, 4.0])
    np.testing.assert_allclose(oracle.interp_table(2.0, rows), [3.0, 3.5, 5.0])
    assert oracle.interp_table(3.0, rows).tolist() == [4.0, 4.0, 4.0]   # t >= last -> last row
    assert oracle.interp_table(99.0, rows).tolist() == [4.0, 4.0, 4.0]
    with pytest.raises(ValueError):                                     # t < first -> reference throws
        oracle.interp_table(-1.0, rows)
