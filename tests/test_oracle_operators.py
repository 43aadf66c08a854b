"""Pins of the oracle's geometry and of its Elasticity / Poisson operators.

What fixes them (the reference binary cannot be built here):
  * test/geometry.cpp:27-111: map, |det J| = 4 and area-scaled normals on the rotated square;
  * discrete divergence theorem and total volume on warped meshes;
  * the assembled path (assemble_volume/skeleton/boundary, an independent restatement of the same
    bilinear form) equals the matrix-free apply;
  * symmetry and definiteness (SIPG with epsilon = -1, Elasticity.h:246; MAT_SYMMETRIC,
    PetscDGMatrix.cpp:76);
  * rigid-body / constant null space with natural BCs on all faces (natural faces are skipped,
    Elasticity.cpp:838-840);
  * linear patch test on affine meshes;
  * manufactured-solution convergence orders (test_convergence_static.py:70-93).
"""
import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, cg

import oracle_lib as O
from cases import warp_nodes
from tandem_amd import scenario as S

BC_D, BC_N = O.BC_DIRICHLET, O.BC_NATURAL


def test_rotated_square_geometry():
    """test/geometry.cpp:27-111 (degree-1 map of the 2-triangle GenMesh square onto a rotated [-1,1]^2)."""
    c = s = np.cos(3.14159265359 / 4.0)
    m = O.Mesh.box(2, [1, 1], 0, [0, 0, 0, 0])

    def T(v):
        x, y = 2.0 * v[..., 0] - 1.0, 2.0 * v[..., 1] - 1.0
        return np.stack([c * x - s * y, c * x + s * y], axis=-1)

    nodes = T(m.affine_nodes(1))
    # map: reference vertices go to the transformed element vertices (GID-sorted)
    exp_x = {0: [(-1.0, -1.0), (1.0, -1.0), (-1.0, 1.0)], 1: [(1.0, -1.0), (-1.0, 1.0), (1.0, 1.0)]}
    for el, xs in exp_x.items():
        for k, x in enumerate(xs):
            assert np.allclose(nodes[el, k], [c * x[0] - s * x[1], c * x[0] + s * x[1]], atol=1e-10)
    op = O.Operator(1, m, 1, np.ones((2, 4)), nodes=nodes)
    assert np.allclose(op.array("absDetJ"), 4.0, atol=1e-9)
    expect = {(0, 0): (np.sqrt(2), -np.sqrt(2)), (0, 1): (-np.sqrt(2), -np.sqrt(2)), (0, 2): (0, 2 * np.sqrt(2)),
              (1, 0): (0, -2 * np.sqrt(2)), (1, 1): (np.sqrt(2), np.sqrt(2)), (1, 2): (-np.sqrt(2), np.sqrt(2))}
    n = op.array("normal").reshape(op.F, op.nq, 2)
    for f, (e0, e1, l0, l1, bc) in enumerate(op.facets()):
        assert np.allclose(n[f], expect[(e0, l0)], atol=1e-9)
        if e1 != e0:
            assert np.allclose(-n[f], expect[(e1, l1)], atol=1e-9)


@pytest.mark.parametrize("D,cells,scheme,N", [(3, 2, 1, 4), (3, 2, 0, 3), (2, 3, 0, 6)])
def test_curvilinear_volume_and_closure(D, cells, scheme, N):
    m = O.Mesh.box(D, [cells] * D, scheme, [BC_D] * (2 * D))
    nodes = warp_nodes(m.affine_nodes(N), 0.02)
    op = O.Operator(1, m, N, np.ones((m.num_elements, O.rule(D, 2 * N + 1)[1].size)), nodes=nodes)
    assert op.array("volume").sum() == pytest.approx(1.0, abs=1e-13)
    # sum over the faces of each element of the integrated outward area-scaled normal vanishes
    n = op.array("normal").reshape(op.F, op.nq, D)
    _, w, _ = O.rule(D - 1, 2 * N + 1)
    fl = np.einsum("fqd,q->fd", n, w)
    acc = np.zeros((op.E, D))
    for f, (e0, e1, l0, l1, bc) in enumerate(op.facets()):
        acc[e0] += fl[f]
        if e1 != e0:
            acc[e1] -= fl[f]
    assert np.abs(acc).max() <= 1e-14
    assert op.array("area").sum() > 0


OPS = [
    # (kind, D, cells, scheme, N, bcs, warp)
    (0, 2, 2, 0, 2, [BC_D] * 4, 0.0),
    (0, 3, 2, 1, 2, [BC_D] * 6, 0.02),
    (0, 3, 1, 0, 4, [BC_D, BC_N, BC_N, BC_D, BC_D, BC_N], 0.02),
    (1, 2, 3, 0, 3, [BC_D, BC_N, BC_D, BC_N], 0.02),
    (1, 3, 2, 1, 2, [BC_D] * 6, 0.02),
]


def _op(kind, D, cells, scheme, N, bcs, warp, seed=1):
    m = O.Mesh.box(D, [cells] * D, scheme, bcs)
    nodes = m.affine_nodes(N)
    if warp:
        nodes = warp_nodes(nodes, warp)
    vc, _ = m.coords(N, nodes)
    # correlation length 1: resolved by these coarse meshes, so the P_N projections of the fields
    # stay positive (tandem's penalty uses the minimum of the projected nodal mu, Elasticity.cpp:255-267)
    lam, mu = S.lame_fields(D, seed, length=1.0)
    if kind == 0:
        return O.Operator(0, m, N, lam(vc), mu(vc), nodes=nodes), nodes
    return O.Operator(1, m, N, mu(vc), nodes=nodes), nodes


@pytest.mark.parametrize("case", OPS)
def test_apply_equals_assembled_and_symmetric(case):
    op, _ = _op(*case)
    A = op.assemble()
    x = np.random.default_rng(0).uniform(-1, 1, op.n)
    y = op.apply(x, nthreads=2)
    assert np.linalg.norm(y - A @ x) <= 1e-12 * np.linalg.norm(y)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    if BC_D in case[5]:
        assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


@pytest.mark.parametrize("D,scheme,N", [(3, 1, 2), (3, 0, 4), (2, 0, 3)])
def test_rigid_body_null_space(D, scheme, N):
    m = O.Mesh.box(D, [2] * D, scheme, [BC_N] * (2 * D))
    nodes = warp_nodes(m.affine_nodes(N), 0.02)
    vc, _ = m.coords(N, nodes)
    lam, mu = S.lame_fields(D, 2)
    op = O.Operator(0, m, N, lam(vc), mu(vc), nodes=nodes)
    X = nodes  # nodal basis on the geometry nodes: coefficients of a field are its nodal values
    modes = [np.broadcast_to(e, X.shape) for e in np.eye(D)]
    if D == 3:
        for w in np.eye(3):
            modes.append(np.cross(w, X))
    else:
        modes.append(np.stack([-X[..., 1], X[..., 0]], axis=-1))
    A_scale = np.linalg.norm(op.apply(np.random.default_rng(0).uniform(-1, 1, op.n)))
    for r in modes:
        v = np.ascontiguousarray(np.transpose(r, (0, 2, 1))).ravel()  # (E, Q, Nbf) block layout
        assert np.linalg.norm(op.apply(v)) <= 1e-11 * A_scale * np.linalg.norm(v) / np.sqrt(op.n)
    # Poisson: constants (modal basis: the constant is the first Dubiner mode)
    opp = O.Operator(1, m, N, mu(vc), nodes=nodes)
    v = np.zeros(opp.n)
    v[::opp.bs] = 1.0
    assert np.abs(opp.apply(v)).max() <= 1e-11


def test_linear_patch_test():
    """(A u_lin)_K = 0 on every element without boundary faces (affine mesh, linear u)."""
    D, N = 3, 3
    m = O.Mesh.box(D, [3] * D, 1, [BC_D] * 6)
    nodes = m.affine_nodes(N)
    vc, _ = m.coords(N, nodes)
    # constant coefficients: sigma(u_lin) is constant, so the SIPG residual vanishes in the interior
    op = O.Operator(0, m, N, np.full(vc.shape[:2], 2.0), np.full(vc.shape[:2], 1.0), nodes=nodes)
    G = np.array([[0.3, -0.2, 0.5], [0.1, 0.7, -0.4], [-0.6, 0.2, 0.9]])
    u = nodes @ G.T + np.array([0.1, -0.2, 0.3])
    v = np.ascontiguousarray(np.transpose(u, (0, 2, 1))).ravel()
    y = op.apply(v).reshape(op.E, op.bs)
    fac = op.facets()
    bnd = set(fac[fac[:, 0] == fac[:, 1], 0].tolist())
    interior = [e for e in range(op.E) if e not in bnd]
    assert len(interior) > 0
    assert np.abs(y[interior]).max() <= 1e-11 * np.abs(y).max()


def _solve(op, b):
    A = LinearOperator((op.n, op.n), matvec=lambda v: op.apply(v, nthreads=4))
    x, info = cg(A, b, rtol=1e-12, atol=0.0, maxiter=20000)
    assert info == 0
    return x


def _l2_error(op, x, exact):
    uh, wj = op.eval_volume(x)
    return np.sqrt(np.sum(wj[:, None, :] * (uh - exact) ** 2))


def test_poisson_mms_convergence_p2():
    """c0's problem (K = 1, u = sin(pi x) sin(pi y)) on 4x4x2 .. 16x16x2 P2 triangles:
    L2 order in (N, N+1) = (2, 3) as test_convergence_static.py:70-93 requires."""
    errs = []
    for n in (4, 8, 16):
        m = O.Mesh.box(2, [n, n], 0, [BC_D] * 4)
        vc, fc = m.coords(2)
        op = O.Operator(1, m, 2, np.ones(vc.shape[:2]))
        u, f = S.poisson_sine(vc)
        b = op.rhs(f[:, None, :], np.zeros((op.F, 1, op.nq)))
        errs.append(_l2_error(op, _solve(op, b), u[:, None, :]))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 2.0) and np.all(rates < 3.5), rates


def test_elasticity_mms_convergence():
    """Cosine manufactured solution (examples/elasticity/3d/cosine.lua) with variable log-normal
    Lame fields on warped P2 tets: L2 order approaches N+1 = 3."""
    errs = []
    lam, mu = S.lame_fields(3, 9, length=1.0)
    for n in (2, 3, 4):
        m = O.Mesh.box(3, [n] * 3, 1, [BC_D] * 6)
        nodes = warp_nodes(m.affine_nodes(2), 0.02)
        vc, fc = m.coords(2, nodes)
        op = O.Operator(0, m, 2, lam(vc), mu(vc), nodes=nodes)
        f = S.elasticity_body_force(vc, lam, mu)
        g, _ = S.cosine_solution(fc)
        b = op.rhs(np.ascontiguousarray(f.transpose(0, 2, 1)), np.ascontiguousarray(g.transpose(0, 2, 1)))
        u, _ = S.cosine_solution(vc)
        errs.append(_l2_error(op, _solve(op, b), np.transpose(u, (0, 2, 1))))
    h = 1.0 / np.array([2, 3, 4])
    rates = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(h[:-1] / h[1:])
    assert rates[-1] > 2.4 and rates[-1] < 3.6, rates
