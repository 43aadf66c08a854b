"""Manufactured-solution convergence of the GPU path at P4 (test/test_convergence_static.py:70-93):
the cosine solution of examples/elasticity/3d/cosine.lua:16-45 with variable log-normal Lame fields
on warped Kuhn-cube tet meshes of 3, 4, 6 and 8 cubes per side.  Right-hand side on the GPU
(body force -div sigma(u*) and the Dirichlet data u* on every face), FP64 CG with the element-block
Jacobi preconditioner to a relative residual of 1e-11, errors at the volume points with the
weights W |det J|.  As in the reference, the convergence order is the slope of a least-squares fit
of log(error) against log(h); bars: L2 order in (N, N+1) = (4, 5), H1-semi order in (N-1, N) =
(3, 4).
"""
import numpy as np
import pytest

import tandem_amd as T
from tandem_amd import scenario as S

pytestmark = pytest.mark.gpu

N = 4


def _errors(cells, lam, mu):
    import torch

    mesh = T.Mesh.box(3, cells, scheme=1, bc=T.BC_DIRICHLET)
    nodes = S.warp(mesh.affine_nodes(N))
    op = T.DGOperator(mesh, T.ELASTICITY, N, nodes)
    X = op.volume_points()                                   # (E, Nq, 3)
    op.set_material(lam(X), mu(X))
    f = S.elasticity_body_force(X, lam, mu)                  # (E, Nq, 3)
    g, _ = S.cosine_solution(op.facet_points())              # (F, nq, 3)
    fd = torch.from_numpy(np.ascontiguousarray(f.transpose(0, 2, 1)).ravel()).cuda()
    gd = torch.from_numpy(np.ascontiguousarray(g.transpose(0, 2, 1)).ravel()).cuda()
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.rhs_device(b, fd, gd)
    x = torch.zeros_like(b)
    r = op.solve(b, x, rtol=1e-11, mode="fp64", precond="block-jacobi", max_iter=100000)
    assert r["converged"], r
    U = x.view(op.E, 3, op.Nbf).cpu().numpy()                # (E, 3, Nbf)
    EQ = op.host_array("E_Q").reshape(op.Nq, op.Nbf)          # [q][k]
    DQ = op.host_array("Dxi_Q").reshape(op.Nq, 3, op.Nbf)     # [q][e][k]
    JI = op.host_array("JInv").reshape(op.E, op.Nq, 3, 3)     # [.., s, e]: d xi_e / d x_s
    wj = op.host_array("absDetJ").reshape(op.E, op.Nq) * op.host_array("volWeights")[None, :]
    uh = np.einsum("erk,qk->eqr", U, EQ)
    gref = np.einsum("erk,qik->eqri", U, DQ)                 # d u_r / d xi_i
    gh = np.einsum("eqri,eqsi->eqrs", gref, JI)              # d u_r / d x_s
    ue, Ge = S.cosine_solution(X)
    l2 = np.sqrt(np.sum(wj[..., None] * (uh - ue) ** 2))
    h1 = np.sqrt(np.sum(wj[..., None, None] * (gh - Ge) ** 2))
    print("cells %d: %d DoF, %d iterations, L2 %.4e, H1-semi %.4e" % (cells, op.n, r["iterations"], l2, h1),
          flush=True)
    return l2, h1


def test_p4_manufactured_solution_convergence_orders(cuda):
    lam, mu = S.lame_fields(3, 9, length=1.0)
    cells = np.array([3, 4, 6, 8])
    errs = np.array([_errors(int(c), lam, mu) for c in cells])
    h = 1.0 / cells
    l2_order = np.polyfit(np.log(h), np.log(errs[:, 0]), 1)[0]
    h1_order = np.polyfit(np.log(h), np.log(errs[:, 1]), 1)[0]
    print("orders: L2 %.3f  H1-semi %.3f" % (l2_order, h1_order), flush=True)
    assert N < l2_order < N + 1
    assert N - 1 < h1_order < N
