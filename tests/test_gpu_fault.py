"""Fault adapter and Green's-function columns on the GPU (through the C ABI) against the oracle.

Bar (FP64): the slip at the facet points of random fault states and the fault-basis traction of
random solutions with facet data equal the oracle's (Adapter<Elasticity>::slip / ::traction on the
oracle's facet traction) to 1e-12 relative; Green's-function columns (unit fault states -> slip ->
right-hand side -> solve -> fault traction) equal the oracle's with dense solves to 1e-8 (solver
tolerance 1e-12).  The fault is bent by the geometry map, so normals and normal lengths vary.
"""
import numpy as np
import pytest

import oracle_lib as O
import tandem_amd as T
from cases import fault_plane_mesh, rel_l2

pytestmark = pytest.mark.gpu

CASES = [(2, 4, 0, 4), (3, 2, 1, 2)]


def _setup(D, n, scheme, N):
    bcs = [5, 1, 5, 5, 5, 1][:2 * D]
    pm, om = fault_plane_mesh(D, n, scheme, bc=bcs)
    X = pm.affine_nodes(N)
    nodes = X.copy()
    bend = 0.04 * np.sin(np.pi * X[..., 1]) * (np.sin(np.pi * X[..., 2]) if D == 3 else 1.0)
    nodes[..., 0] += bend * np.sin(np.pi * X[..., 0])
    nodes = np.ascontiguousarray(nodes)
    op = T.DGOperator(pm, T.ELASTICITY, N, nodes)
    Xq = op.volume_points()
    lam = 2.0 + 0.3 * np.sin(2 * Xq[..., 0]) * np.cos(Xq[..., 1])
    mu = 1.0 + 0.2 * np.cos(Xq[..., 0] + 2 * Xq[..., 1])
    op.set_material(lam, mu)
    oo = O.Operator(0, om, N, lam, mu, nodes=nodes)
    up = np.array([0.0, 1.0] if D == 2 else [0.0, 0.0, 1.0])
    refn = np.array([1.0, 0.0] if D == 2 else [1.0, 0.0, 0.0])
    nf, nbf = op.fault_setup(up, refn)
    fac, _, _, _ = oo.fault_setup(up, refn)
    assert nf == len(fac) and nbf == oo.fault_nbf and nf > 0
    return op, oo, fac


@pytest.mark.parametrize("case", CASES, ids=["D%d-n%d-s%d-P%d" % c for c in CASES])
def test_fault_slip_and_traction_match_oracle(cuda, case):
    import torch

    D = case[0]
    op, oo, fac = _setup(*case)
    nf, nbf, nq = op.nfault, op.fault_nbf, op.nq
    print("case", case, "fault facets", nf, "nbf", nbf, flush=True)
    rng = np.random.default_rng(51)
    nc = 3
    S = rng.uniform(-1, 1, (nc, nf * nbf * (D - 1)))
    F = op.fault_slip(torch.from_numpy(S).to(cuda)).cpu().numpy().reshape(nc, op.F, D, nq)
    for c in range(nc):
        ref = np.zeros((op.F, D, nq))
        for m in range(nf):
            s = S[c, m * nbf * (D - 1):(m + 1) * nbf * (D - 1)].reshape(D - 1, nbf).T
            ref[fac[m]] = oo.fault_slip(m, s)
        assert np.abs(F[c] - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    Xs = rng.uniform(-1, 1, (2, op.n))
    G = rng.uniform(-1, 1, (2, op.F * D * nq))
    Tg = op.fault_traction(torch.from_numpy(Xs).to(cuda), torch.from_numpy(G).to(cuda)).cpu().numpy()
    for c in range(2):
        ref = np.concatenate([oo.fault_traction(m, oo.traction(fac[m], Xs[c], G[c].reshape(op.F, D, nq)[fac[m]]))
                              .T.ravel() for m in range(nf)])
        err = rel_l2(Tg[c], ref)
        print("  fault traction rel L2 %.3e" % err, flush=True)
        assert err <= 1e-12
    # without facet data
    Tn = op.fault_traction(torch.from_numpy(Xs[0]).to(cuda)).cpu().numpy()
    ref = np.concatenate([oo.fault_traction(m, oo.traction(fac[m], Xs[0], None)).T.ravel() for m in range(nf)])
    assert rel_l2(Tn, ref) <= 1e-12


def test_green_function_columns_match_dense_oracle(cuda):
    """Unit fault states -> facet slip -> batched right-hand side -> lockstep CG -> fault traction,
    all on the GPU, against the oracle's slip, right-hand side, dense solve, traction and adapter."""
    import torch

    D, N = 2, 4
    op, oo, fac = _setup(D, 4, 0, N)
    nf, nbf, nq = op.nfault, op.fault_nbf, op.nq
    cols = [0, 1, nbf, nf * nbf - 1]
    S = np.zeros((len(cols), nf * nbf * (D - 1)))
    for i, c in enumerate(cols):
        S[i, c] = 1.0
    Sd = torch.from_numpy(S).to(cuda)
    F = op.fault_slip(Sd)
    B = torch.empty((len(cols), op.n), dtype=torch.float64, device=cuda)
    op.rhs_device(B, None, F)
    Xd = torch.zeros_like(B)
    r = op.solve_multi(B, Xd, rtol=1e-12, mode="fp64")
    assert r["converged"]
    Tg = op.fault_traction(Xd, F).cpu().numpy()
    A = oo.assemble()
    Fh = F.cpu().numpy()
    for i in range(len(cols)):
        fq = Fh[i].reshape(op.F, D, nq)
        b = oo.rhs(None, fq)
        x = np.linalg.solve(A, b)
        ref = np.concatenate([oo.fault_traction(m, oo.traction(fac[m], x, fq[fac[m]])).T.ravel() for m in range(nf)])
        err = rel_l2(Tg[i], ref)
        print("column", cols[i], "rel L2 %.3e" % err, flush=True)
        assert err <= 1e-8
