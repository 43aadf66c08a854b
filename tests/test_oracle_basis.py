"""Pins of the oracle's polynomials, nodes and quadrature against the reference's own unit tests
(test/basis.cpp, test/quadrules.cpp, test/form.cpp) and against closed forms.

Golden vectors: tests/golden/reference_golden.json (written by tests/golden/extract_golden.py from
the reference test sources).  Closed-form Dubiner polynomials: test/basis.cpp:44-206.
"""
import json
import os
from math import comb, factorial, sqrt

import numpy as np
import pytest

import oracle_lib as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))


@pytest.mark.parametrize("dim,p", [(d, p) for d in (2, 3) for p in (1, 2, 3, 4)])
def test_warp_and_blend_nodes_golden(dim, p):
    """test/basis.cpp:264-407: nodes agree with the stored values (doctest Approx ~ 1e-5 there;
    here to 1e-14)."""
    ref = np.array(GOLD["wab_%dd_p%d" % (dim, p)])
    got = O.wab_nodes(dim, p)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-14


def test_warp_and_blend_invariants():
    """Higher degrees (no stored values): nodes lie in the simplex, include the vertices, are
    distinct, and the nodal Vandermonde is well conditioned (Lebesgue-type check)."""
    for dim, N in ((2, 6), (3, 4), (2, 5)):
        X = O.wab_nodes(dim, N)
        assert X.shape == (comb(N + dim, dim), dim)
        assert X.min() >= -1e-14 and X.sum(axis=1).max() <= 1 + 1e-14
        for v in np.vstack([np.zeros(dim), np.eye(dim)]):
            assert np.abs(X - v).sum(axis=1).min() <= 1e-14
        d = np.linalg.norm(X[:, None] - X[None], axis=-1) + np.eye(len(X))
        assert d.min() > 1e-3


TRI = {
    (0, 0): lambda x, y: 1.0 + 0 * x,
    (1, 0): lambda x, y: -1.0 + 2.0 * x + y,
    (0, 1): lambda x, y: -1.0 + 3.0 * y,
    (2, 0): lambda x, y: 1.0 - 6.0 * x + 6.0 * x * x - 2.0 * y + 6.0 * x * y + y * y,
    (1, 1): lambda x, y: 1.0 - 2.0 * x - 6.0 * y + 10.0 * x * y + 5.0 * y * y,
    (0, 2): lambda x, y: 1.0 - 8.0 * y + 10.0 * y * y,
    (3, 0): lambda x, y: -1.0 + 12.0 * x - 30.0 * x * x + 20.0 * x ** 3 + 3.0 * y - 24.0 * x * y
    + 30.0 * x * x * y - 3.0 * y * y + 12.0 * x * y * y + y ** 3,
    (0, 3): lambda x, y: -1.0 + 15.0 * y - 45.0 * y * y + 35.0 * y ** 3,
}
TET = {
    (0, 0, 0): lambda x, y, z: 1.0 + 0 * x,
    (1, 0, 0): lambda x, y, z: -1.0 + 2.0 * x + y + z,
    (0, 1, 0): lambda x, y, z: -1.0 + 3.0 * y + z,
    (0, 0, 1): lambda x, y, z: -1.0 + 4.0 * z,
    (2, 0, 0): lambda x, y, z: 1.0 - 6.0 * x + 6.0 * x * x - 2.0 * y + 6.0 * x * y + y * y - 2.0 * z
    + 6.0 * x * z + 2.0 * y * z + z * z,
    (1, 1, 0): lambda x, y, z: 1.0 - 2.0 * x - 6.0 * y + 10.0 * x * y + 5.0 * y * y - 2.0 * z + 2.0 * x * z
    + 6.0 * y * z + z * z,
    (0, 2, 0): lambda x, y, z: 1.0 - 8.0 * y + 10.0 * y * y - 2.0 * z + 8.0 * y * z + z * z,
    (1, 0, 1): lambda x, y, z: 1.0 - 2.0 * x - y - 7.0 * z + 12.0 * x * z + 6.0 * y * z + 6.0 * z * z,
    (0, 1, 1): lambda x, y, z: 1.0 - 3.0 * y - 7.0 * z + 18.0 * y * z + 6.0 * z * z,
    (0, 0, 2): lambda x, y, z: 1.0 - 10.0 * z + 15.0 * z * z,
}


def _num_grad(f, p, h=1e-6):
    p = np.asarray(p, dtype=float)
    return np.array([(f(*(p + h * e)) - f(*(p - h * e))) / (2 * h) for e in np.eye(len(p))])


def test_dubiner_closed_forms():
    """test/basis.cpp:44-206: values and gradients vs the closed-form polynomials."""
    pts2 = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (0.25, 0.25), (0.1, 0.1), (0.1, 0.2), (0.2, 0.1)]
    for idx, f in TRI.items():
        for p in pts2:
            assert O.dubiner(2, idx, p) == pytest.approx(f(*p), abs=1e-13)
            assert np.allclose(O.grad_dubiner(2, idx, p), _num_grad(f, p), atol=1e-7)
    pts3 = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.25, 0.25, 0.25),
            (0.1, 0.1, 0.1), (0.1, 0.2, 0.3), (0.2, 0.1, 0.3)]
    for idx, f in TET.items():
        for p in pts3:
            assert O.dubiner(3, idx, p) == pytest.approx(f(*p), abs=1e-13)
            assert np.allclose(O.grad_dubiner(3, idx, p), _num_grad(f, p), atol=1e-7)


def test_dubiner_orthogonality_and_mass_golden():
    """test/form.cpp:19-57: modal tetrahedron mass matrix is diagonal with the stored diagonal."""
    ref = np.array(GOLD["modal_mass_tet_diag"])
    for N in range(4):
        M = O.ref_mass(3, N, modal=True)
        n = M.shape[0]
        assert np.allclose(np.diag(M), ref[:n], rtol=1e-13, atol=0)
        Minv = O.ref_mass(3, N, modal=True, inverse=True)
        assert np.allclose(np.diag(Minv), 1.0 / ref[:n], rtol=1e-13)
    # orthogonality with an independent exact rule
    pts, w, _ = O.rule(3, 8, family=1)
    E, _ = O.eval_basis(3, 3, True, pts)
    G = (E * w) @ E.T
    assert np.abs(G - np.diag(np.diag(G))).max() <= 1e-14


def test_jacobi_singularity_free_identity():
    """test/basis.cpp:31-40: P_n^{(a,b)}(x/y) y^n computed without division."""
    for n in range(5):
        for a in range(10):
            for b in range(4):
                for x in (0.0, 1.0, 0.25, 0.1, 0.2):
                    lhs = O.lib().oracle_jacobi_p(n, a, b, x / 2.0) * 2.0 ** n
                    assert O.lib().oracle_sf_jacobi_p(n, a, b, x, 2.0) == pytest.approx(lhs, rel=1e-12, abs=1e-12)


def test_gauss_jacobi_maple():
    """test/quadrules.cpp:24-37."""
    p, w = O.gauss_jacobi(5, 1, 3)
    g = GOLD["gauss_jacobi_5_1_3"]
    assert np.allclose(p, g["points"], rtol=0, atol=1e-15)
    assert np.allclose(w, g["weights"], rtol=1e-14, atol=0)


def test_lgl_points():
    """test/basis.cpp:210-231."""
    assert np.allclose(O.lgl_points(2), [-1, 1])
    assert np.allclose(O.lgl_points(3), [-1, 0, 1], atol=1e-15)
    assert np.allclose(O.lgl_points(5), [-1, -sqrt(3 / 7), 0, sqrt(3 / 7), 1], atol=1e-15)


def test_collapsed_triangle_rule_maple():
    """test/quadrules.cpp:38-52."""
    pts, w, _ = O.rule(2, 3, family=1)
    g = GOLD["tensor_triangle_order3"]
    assert np.allclose(pts, g["points"], atol=1e-15)
    assert np.allclose(w, g["weights"], atol=1e-15)


def _monomial_exact(D, j):
    j = sorted(j)
    ref = 1.0 / np.prod(np.arange(j[-1] + 1, sum(j) + D + 1, dtype=float))
    for d in range(D - 1):
        ref *= factorial(j[d])
    return ref


@pytest.mark.parametrize("D,family,orders", [
    (1, 1, range(1, 22, 2)), (2, 1, range(1, 22, 2)), (3, 1, range(1, 22, 2)),
    (2, 2, range(1, 21)), (3, 2, range(1, 11)), (3, 3, range(1, 14))])
def test_rule_monomial_exactness(D, family, orders):
    """test/quadrules.cpp:54-120: every monomial of degree <= order integrates to
    j_1!...j_D!/(|j|+D)!.  WV/JS tables are the extracted published data (orders JS <= 13)."""
    for o in orders:
        r = O.rule(D, o, family=family)
        assert r is not None
        pts, w, order = r
        assert np.all(w > 0)
        for j in O.all_integer_sums(D, order):
            val = np.sum(w * np.prod(pts ** j, axis=1))
            assert val == pytest.approx(_monomial_exact(D, j), rel=1e-12, abs=1e-15)


def test_autorule_sizes():
    """test/quadrules.cpp:122-133 (3D orders 20/21 need JS tables beyond order 13, which this
    project does not extract; the 3D order-9 rule of the P4 configs is JS2020's 57-point rule)."""
    for D, o, n in GOLD["autorule_sizes"]:
        if D == 3 and o >= 14:
            continue
        assert O.rule(D, o)[1].size == n
    assert O.rule(3, 9)[1].size == 57 and O.rule(2, 9)[1].size == 19
    assert O.rule(2, 5)[1].size == 7 and O.rule(2, 13)[1].size == 37


def test_nodal_projection_of_cubic():
    """test/form.cpp:60-90: L2 projection of a cubic onto the nodal P3 line element is exact."""
    f = lambda x: x ** 3 - 4.0 * x * x + 5.0 * x - 1.0
    pts, w, _ = O.rule(1, 6)
    E, _ = O.eval_basis(1, 3, False, pts)
    from math import comb as _c  # noqa: F401
    M = (E * w) @ E.T
    F = np.linalg.solve(M, E @ (w * f(pts[:, 0])))
    nodes = O.wab_nodes(1, 3)[:, 0]
    assert np.allclose(F, f(nodes), atol=1e-13)
    assert np.allclose(E.T @ F, f(pts[:, 0]), atol=1e-13)


def test_integer_sum_order():
    """Combinatorics.h:46-110: the enumeration tandem uses for basis and node indices."""
    assert O.all_integer_sums(2, 2).tolist() == [[0, 0], [1, 0], [0, 1], [2, 0], [1, 1], [0, 2]]
    t = O.all_integer_sums(3, 4)
    assert len(t) == 35 and t[:5].tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 0]]
    assert all(t.sum(axis=1)[i] <= t.sum(axis=1)[i + 1] for i in range(len(t) - 1))
