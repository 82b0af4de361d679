"""The oracle's free-form deformation (femPointGrid, src/femPointGrid.cpp),
pinned by what the mathematics fixes: Bernstein partition of unity and
linear precision, box coordinates that rebuild the point (X = X0 + sS +
tT + uU, oblique axes included), zero control displacements -> zero
morph, and the reference's storage of box coordinates for points outside
the box (:193-197)."""
import numpy as np
import pytest

from oracle import oracle

S0, T0, U0 = np.array([2.0, 0.0, 0.0]), np.array([0.0, 3.0, 0.0]), np.array([0.0, 0.0, 4.0])


def _grid_index(n, i, j, k):
    return i + n[0] * (j + n[1] * k)


def test_zero_displacements_and_outside_points():
    n = (3, 4, 3)
    pts = np.array([[1.0, 1.5, 2.0], [0.5, 0.1, 3.9], [3.0, 1.0, 1.0], [-0.5, 1.0, 1.0]])
    out = oracle.ffd_morph([0, 0, 0], n, S0, T0, U0, [], np.zeros((0, 3)), pts)
    assert np.array_equal(out[:2], np.zeros((2, 3)))
    # outside: the box coordinates are stored (reference behaviour)
    np.testing.assert_allclose(out[2], [1.5, 1.0 / 3.0, 0.25], rtol=1e-15)
    np.testing.assert_allclose(out[3], [-0.25, 1.0 / 3.0, 0.25], rtol=1e-15)


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3), (5, 4, 6), (13, 3, 2)])
def test_partition_of_unity(n):
    """Every control point moved by the same (a, b, c) along S, T, U: every
    inside point moves by a S/|S| + b T/|T| + c U/|U|."""
    tot = n[0] * n[1] * n[2]
    d = np.array([0.3, -0.2, 0.05])
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.0, 1.0, (50, 3)) * [2.0, 3.0, 4.0]
    out = oracle.ffd_morph([0, 0, 0], n, S0, T0, U0, np.arange(tot), np.tile(d, (tot, 1)), pts)
    np.testing.assert_allclose(out, np.tile(d, (50, 1)), rtol=1e-12, atol=1e-14)


def test_linear_precision_oblique_box():
    """d_ijk = (i / (nx - 1)) v along S on an oblique box: the morph is s v,
    with s the point's S coordinate (sum_i (i/n) B_i^n(s) = s)."""
    n = (4, 3, 5)
    S, T, U = np.array([2.0, 0.5, 0.0]), np.array([0.2, 3.0, 0.1]), np.array([0.0, -0.4, 4.0])
    X0 = np.array([0.5, -1.0, 2.0])
    rng = np.random.default_rng(2)
    stu = rng.uniform(0.0, 1.0, (40, 3))
    pts = X0 + stu[:, :1] * S + stu[:, 1:2] * T + stu[:, 2:3] * U
    v = 0.7
    nodes, vals = [], []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                nodes.append(_grid_index(n, i, j, k))
                vals.append([v * i / (n[0] - 1), 0.0, 0.0])
    out = oracle.ffd_morph(X0, n, S, T, U, nodes, vals, pts)
    expect = stu[:, :1] * v * (S / np.linalg.norm(S))
    np.testing.assert_allclose(out, expect, rtol=1e-12, atol=1e-14)
    # box coordinates rebuild the points (outside points report them)
    far = X0 + 1.5 * S + 0.25 * T - 0.5 * U
    o2 = oracle.ffd_morph(X0, n, S, T, U, [], np.zeros((0, 3)), far[None])
    np.testing.assert_allclose(o2[0], [1.5, 0.25, -0.5], rtol=1e-13, atol=1e-14)


def test_single_control_point_is_the_tensor_bernstein_basis():
    n = (3, 3, 3)
    i, j, k = 1, 2, 0
    pts = np.array([[0.8, 1.2, 1.0], [1.9, 2.5, 0.3]])
    out = oracle.ffd_morph([0, 0, 0], n, S0, T0, U0, [_grid_index(n, i, j, k)], [[1.0, 0.0, 0.0]], pts)
    s, t, u = pts[:, 0] / 2.0, pts[:, 1] / 3.0, pts[:, 2] / 4.0
    b = (2 * s * (1 - s)) * (t * t) * ((1 - u) ** 2)  # C(2,1) s (1-s), C(2,2) t^2, C(2,0) (1-u)^2
    np.testing.assert_allclose(out[:, 0], b, rtol=1e-14)
    assert not out[:, 1:].any()


def test_bad_control_point_index():
    with pytest.raises(ValueError):
        oracle.ffd_morph([0, 0, 0], (2, 2, 2), S0, T0, U0, [8], [[1.0, 0, 0]], np.zeros((1, 3)))
