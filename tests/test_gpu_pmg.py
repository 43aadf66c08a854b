"""p-multigrid over polynomial degree (through the C ABI): the coarse operator with the fine facet
penalty is the Galerkin product P^T A P, the V-cycle is symmetric and positive, the eigenvalue
estimate agrees with power iteration, and the flexible-CG solve reaches 1e-10 with the block-Jacobi
CG solution in far fewer iterations.

The reference forms its coarse matrices as P^T A P (app/common/PetscLinearSolver.cpp:83-103) with
the interpolation of NodalInterpolation::assemble (app/localoperator/NodalInterpolation.cpp:8-38).
On affine meshes with constant material every quadrature in the coarse operator is exact, so the
matrix-free coarse operator must reproduce P^T A P to rounding.
"""
import numpy as np
import pytest

import tandem_amd as T
from tandem_amd import scenario as S

pytestmark = pytest.mark.gpu


def make(mesh, degree, lam, mu):
    op = T.DGOperator(mesh, T.ELASTICITY, degree)
    X = op.volume_points()
    op.set_material(lam(X), mu(X))
    return op


def const(v):
    return lambda X: np.full(X.shape[:-1], v)


@pytest.mark.parametrize("dim,fine,coarse", [(2, 6, 3), (2, 6, 1), (3, 4, 2), (3, 3, 1)])
def test_coarse_operator_is_galerkin_product(cuda, dim, fine, coarse):
    import torch

    mesh = T.Mesh.box(dim, 4 if dim == 2 else 2, scheme=1 if dim == 3 else 0, bc=T.BC_DIRICHLET)
    f = make(mesh, fine, const(2.0), const(1.0))
    c = make(mesh, coarse, const(2.0), const(1.0))
    T.PMultigrid([f, c])  # the coarse level takes the fine penalty
    I = torch.from_numpy(T.interpolation_matrix(dim, fine, coarse)).to(cuda)
    xc = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, c.n)).to(cuda)
    yc = torch.zeros_like(xc)
    c.apply(xc, yc)
    xf = T.interpolate(I, xc, dim)
    yf = torch.zeros(f.n, dtype=torch.float64, device=cuda)
    f.apply(xf, yf)
    torch.cuda.synchronize()
    ref = T.interpolate(I, yf, dim, transpose=True)
    torch.cuda.synchronize()
    err = (torch.linalg.norm(yc - ref) / torch.linalg.norm(ref)).item()
    print("D%d P%d->P%d  |A_c x - P^T A P x| / |P^T A P x| = %.3e" % (dim, fine, coarse, err), flush=True)
    assert err <= 1e-11


def hierarchy(mesh, degrees, seed, **kw):
    lam, mu = S.lame_fields(mesh.dim, seed)
    ops = [make(mesh, p, lam, mu) for p in degrees]
    return ops, T.PMultigrid(ops, **kw)


def test_vcycle_symmetric_positive_and_eigen_estimate(cuda):
    import torch

    mesh = T.Mesh.box(2, 6, bc=T.BC_DIRICHLET)
    ops, M = hierarchy(mesh, (6, 3, 1), 11, coarse_rtol=1e-14, coarse_max_iter=5000, coarse_fp32=0,
                       vcycle_fp32=0)
    lmax = M.setup()
    print("lambda_max per level", lmax, flush=True)
    n = ops[0].n
    rng = np.random.default_rng(5)
    X = torch.from_numpy(rng.uniform(-1, 1, (2, n))).to(cuda)
    Z = torch.zeros_like(X)
    M.vcycle(X, Z)
    torch.cuda.synchronize()
    a = torch.dot(X[0], Z[1]).item()
    b = torch.dot(Z[0], X[1]).item()
    sym = abs(a - b) / (torch.linalg.norm(X[0]) * torch.linalg.norm(Z[1])).item()
    pos = [torch.dot(X[i], Z[i]).item() for i in range(2)]
    print("x.My %.12e  Mx.y %.12e  rel %.3e  x.Mx %s" % (a, b, sym, pos), flush=True)
    assert sym <= 1e-9
    assert min(pos) > 0.0
    # the FP32 V-cycle (FP32 smoothing, transfers and coarse solve) is the same map to FP32 accuracy
    M32 = T.PMultigrid(ops, coarse_rtol=1e-14, coarse_max_iter=5000, coarse_fp32=1, vcycle_fp32=1)
    Z32 = torch.zeros_like(X)
    M32.vcycle(X, Z32)
    torch.cuda.synchronize()
    d32 = (torch.linalg.norm(Z32 - Z) / torch.linalg.norm(Z)).item()
    print("FP32 V-cycle vs FP64: %.3e" % d32, flush=True)
    assert d32 <= 1e-4
    # the coarsest level ran on its assembled block rows (the default where the block size has the
    # register kernel); the matrix-free coarse operator gives the same V-cycle to the coarse tolerance
    ops[-1].set_variant("auto")
    M0 = T.PMultigrid(ops, coarse_rtol=1e-14, coarse_max_iter=5000, coarse_fp32=0, vcycle_fp32=0,
                      coarse_assembled=0)
    Z0 = torch.zeros_like(X)
    M0.vcycle(X, Z0)
    torch.cuda.synchronize()
    d0 = (torch.linalg.norm(Z0 - Z) / torch.linalg.norm(Z)).item()
    print("matrix-free coarse level vs assembled: %.3e" % d0, flush=True)
    assert d0 <= 1e-10
    # power iteration on D^-1 A of the fine level: the Lanczos estimate is a lower bound of the
    # largest eigenvalue, as is every power iterate
    op = ops[0]
    v = torch.from_numpy(rng.uniform(-1, 1, n)).to(cuda)
    w = torch.zeros_like(v)
    z = torch.zeros_like(v)
    lam = 0.0
    for _ in range(300):
        v /= torch.linalg.norm(v)
        op.apply(v, w)
        op.block_jacobi_apply(w, z)
        torch.cuda.synchronize()
        lam = torch.dot(v, z).item()
        v, z = z, v
    print("power iteration %.6f  Lanczos %.6f" % (lam, lmax[0]), flush=True)
    assert abs(lmax[0] - lam) <= 0.05 * lam


@pytest.mark.parametrize("dim,degrees,n,fp32", [(2, (6, 3, 1), 8, 1), (2, (6, 1), 8, 0), (3, (4, 2, 1), 3, 1)])
def test_pmg_solve_matches_block_jacobi_cg(cuda, dim, degrees, n, fp32):
    import torch

    mesh = T.Mesh.box(dim, n, scheme=1 if dim == 3 else 0, bc=T.BC_DIRICHLET)
    ops, M = hierarchy(mesh, degrees, 7, vcycle_fp32=fp32, coarse_fp32=fp32)
    op = ops[0]
    nc = 3
    B = torch.from_numpy(np.random.default_rng(8).uniform(-1, 1, (nc, op.n))).to(cuda)
    X = torch.zeros_like(B)
    r = M.solve_multi(B, X, rtol=1e-10, max_iter=300)
    print("p-MG", degrees, "fp32" if fp32 else "fp64",
          {k: r[k] for k in ("iterations", "outer_iterations", "rel_residual", "seconds")}, flush=True)
    X2 = torch.zeros_like(B)
    r2 = op.solve_multi(B, X2, rtol=1e-13, mode="fp64", precond="block-jacobi")
    print("block-Jacobi CG", {k: r2[k] for k in ("iterations", "rel_residual")}, flush=True)
    assert r["converged"] == 1 and r["rel_residual"] <= 1e-10
    assert r2["rel_residual"] <= 1e-11
    for c in range(nc):
        err = (torch.linalg.norm(X[c] - X2[c]) / torch.linalg.norm(X2[c])).item()
        print("  column %d: |x_pmg - x_bj| / |x_bj| = %.3e" % (c, err), flush=True)
        assert err <= 1e-6
    assert r["iterations"] * 5 <= r2["iterations"]


@pytest.mark.parametrize("dim,cells,degrees,fp32", [(2, 24, (4, 1), 1), (2, 12, (4, 2), 1), (3, 4, (2, 1), 1),
                                                   (2, 24, (4, 1), 0)],
                         ids=["D2-P4-P1", "D2-P4-P2", "D3-P2-P1", "D2-P4-P1-fp64"])
def test_amg_coarse_preconditioner(cuda, dim, cells, degrees, fp32):
    """The coarse CG preconditioned by the smoothed-aggregation AMG (the default for elasticity P1 / 2D
    P2 coarse levels solved in FP32; the reference's coarse preconditioner is GAMG) against block
    Jacobi: the same solution to the solver tolerance, and a fraction of the coarse iterations.  The
    three hierarchies cover the AMG's block shapes (element blocks of 6, 12 and 12 with 3, 3 and 6
    rigid-body modes); the last runs the V-cycle and the coarse CG (with its AMG) in FP64."""
    import torch

    mesh = T.Mesh.box(dim, cells, scheme=1 if dim == 3 else 0, bc=T.BC_DIRICHLET)
    B = None
    out = {}
    for amg in (0, 1):
        ops, M = hierarchy(mesh, degrees, 7, coarse_amg=amg, coarse_fp32=fp32, vcycle_fp32=fp32)
        if B is None:
            B = torch.from_numpy(np.random.default_rng(9).uniform(-1, 1, (4, ops[0].n))).to(cuda)
        X = torch.zeros_like(B)
        r = M.solve_multi(B, X, rtol=1e-10, max_iter=300)
        print("coarse_amg", amg, {k: r[k] for k in ("iterations", "outer_iterations", "rel_residual", "seconds")},
              flush=True)
        assert r["converged"] == 1 and r["rel_residual"] <= 1e-10
        out[amg] = (X.clone(), r["outer_iterations"], r["iterations"])
    err = (torch.linalg.norm(out[1][0] - out[0][0]) / torch.linalg.norm(out[0][0])).item()
    print("AMG vs block-Jacobi coarse CG: solution rel diff %.3e" % err, flush=True)
    assert err <= 1e-7
    assert out[1][1] * 3 <= out[0][1]
