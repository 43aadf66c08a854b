"""GPU right-hand side and facet traction (through the C ABI) against the CPU FP64 oracle.

Bar: FP64 relative L2 <= 1e-12 (rhs) and max abs <= 1e-11 relative to the largest entry (traction)
on meshes with an interior fault plane, Dirichlet, Fault and natural box faces, warped geometry and
variable coefficients; the batched rhs equals the single-column rhs bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O
import tandem_amd as T
from cases import fault_plane_mesh, rel_l2, smooth_coefficients, warp_nodes

pytestmark = pytest.mark.gpu

CASES = [(3, 2, 1, 4), (3, 2, 0, 2), (2, 4, 0, 6), (2, 4, 0, 3)]


def _setup(D, n, scheme, N):
    bcs = [5, 3, 1, 5, 5, 3][:2 * D]
    pm, om = fault_plane_mesh(D, n, scheme, bc=bcs)
    nodes = np.ascontiguousarray(warp_nodes(pm.affine_nodes(N), 0.02))
    op = T.DGOperator(pm, T.ELASTICITY, N, nodes)
    lam, mu = smooth_coefficients(op.volume_points(), 5)
    op.set_material(lam, mu)
    oo = O.Operator(0, om, N, lam, mu, nodes=nodes)
    return op, oo


@pytest.mark.parametrize("case", CASES, ids=["D%d-n%d-s%d-P%d" % c for c in CASES])
def test_rhs_device_matches_oracle(cuda, case):
    import torch

    D, n, scheme, N = case
    op, oo = _setup(D, n, scheme, N)
    rng = np.random.default_rng(8)
    f = rng.uniform(-1, 1, (op.E, D, op.Nq))
    g = rng.uniform(-1, 1, (op.F, D, op.nq))
    ref = oo.rhs(f, g)
    print("case", case, "n", op.n, flush=True)
    b = torch.full((op.n,), float("nan"), dtype=torch.float64, device=cuda)
    op.rhs_device(b, torch.from_numpy(f.ravel()).to(cuda), torch.from_numpy(g.ravel()).to(cuda))
    err = rel_l2(b.cpu().numpy(), ref)
    print("  rhs rel L2 %.3e" % err, flush=True)
    assert err <= 1e-12
    # force only and facet data only
    op.rhs_device(b, torch.from_numpy(f.ravel()).to(cuda), None)
    assert rel_l2(b.cpu().numpy(), oo.rhs(f, None)) <= 1e-12
    op.rhs_device(b, None, torch.from_numpy(g.ravel()).to(cuda))
    assert rel_l2(b.cpu().numpy(), oo.rhs(None, g)) <= 1e-12


def test_rhs_device_batched_columns(cuda):
    import torch

    op, oo = _setup(3, 2, 1, 4)
    rng = np.random.default_rng(9)
    nc = 3
    F = torch.from_numpy(rng.uniform(-1, 1, (nc, op.E * 3 * op.Nq + 5))).to(cuda)
    G = torch.from_numpy(rng.uniform(-1, 1, (nc, op.F * 3 * op.nq + 3))).to(cuda)
    B = torch.zeros((nc, op.n + 7), dtype=torch.float64, device=cuda)
    op.rhs_device(B, F, G)
    for c in range(nc):
        b = torch.zeros(op.n, dtype=torch.float64, device=cuda)
        op.rhs_device(b, F[c, :op.E * 3 * op.Nq].contiguous(), G[c, :op.F * 3 * op.nq].contiguous())
        assert torch.equal(b, B[c, :op.n])


@pytest.mark.parametrize("case", CASES, ids=["D%d-n%d-s%d-P%d" % c for c in CASES])
def test_traction_matches_oracle(cuda, case):
    import torch

    D, n, scheme, N = case
    op, oo = _setup(D, n, scheme, N)
    rng = np.random.default_rng(10)
    x = rng.uniform(-1, 1, op.n)
    g = rng.uniform(-1, 1, (op.F, D, op.nq))
    t = op.traction(torch.from_numpy(x).to(cuda), torch.from_numpy(g.ravel()).to(cuda)).cpu().numpy()
    ref = np.stack([oo.traction(f, x, g[f]) for f in range(op.F)])
    err = np.abs(t - ref).max() / np.abs(ref).max()
    print("case", case, "traction max rel %.3e" % err, flush=True)
    assert err <= 1e-11
