"""GPU elastic-wave pieces and p-multigrid transfer (through the C ABI) against the CPU FP64 oracle.

Bar: FP64 relative L2 <= 1e-12 for the weight-adjusted inverse mass and the wave right-hand side
(with and without facet data, on meshes with an interior fault plane, warped geometry and variable
coefficients); prolongation / restriction equal I x / I^T x of the oracle's interpolation matrix to
1e-13, and prolongation reproduces a degree-`from` field exactly at the finer nodes.
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
def test_inverse_mass_matches_oracle(cuda, case):
    import torch

    op, oo = _setup(*case)
    print("case", case, "n", op.n, flush=True)
    x = np.random.default_rng(21).uniform(-1, 1, op.n)
    y = op.inverse_mass(torch.from_numpy(x).to(cuda)).cpu().numpy()
    err = rel_l2(y, oo.inverse_mass(x))
    print("  inverse mass rel L2 %.3e" % err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("case", CASES, ids=["D%d-n%d-s%d-P%d" % c for c in CASES])
def test_wave_rhs_matches_oracle(cuda, case):
    import torch

    D = case[0]
    op, oo = _setup(*case)
    print("case", case, "n", op.n, flush=True)
    rng = np.random.default_rng(22)
    x = rng.uniform(-1, 1, op.n)
    g = rng.uniform(-1, 1, (op.F, D, op.nq))
    xd = torch.from_numpy(x).to(cuda)
    y = op.wave_rhs(xd, torch.from_numpy(g.ravel()).to(cuda)).cpu().numpy()
    err = rel_l2(y, oo.wave_rhs(x, g))
    print("  wave rhs rel L2 %.3e" % err, flush=True)
    assert err <= 1e-12
    y0 = op.wave_rhs(xd).cpu().numpy()
    assert rel_l2(y0, oo.wave_rhs(x, None)) <= 1e-12


@pytest.mark.parametrize("D,to,frm", [(3, 4, 2), (3, 2, 1), (2, 6, 3), (2, 3, 2)])
def test_interpolate_matches_oracle(cuda, D, to, frm):
    import torch

    I = O.nodal_interpolation(D, to, frm)
    nt, nf = I.shape
    E, Q = 37, D
    rng = np.random.default_rng(23)
    Id = torch.from_numpy(np.ascontiguousarray(T.interpolation_matrix(D, to, frm))).to(cuda)
    x = rng.uniform(-1, 1, (E, Q, nf))
    y = T.interpolate(Id, torch.from_numpy(x.ravel()).to(cuda), Q).cpu().numpy().reshape(E, Q, nt)
    ref = np.einsum("kl,epl->epk", I, x)
    assert np.abs(y - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    r = rng.uniform(-1, 1, (E, Q, nt))
    z = T.interpolate(Id, torch.from_numpy(r.ravel()).to(cuda), Q, transpose=True).cpu().numpy().reshape(E, Q, nf)
    refz = np.einsum("kl,epk->epl", I, r)
    assert np.abs(z - refz).max() <= 1e-13 * max(1.0, np.abs(refz).max())
    # a degree-frm polynomial given by its coarse nodal values is reproduced at the fine nodes
    cn, fn = T.reference_nodes(D, frm), T.reference_nodes(D, to)

    def poly(p):  # total degree frm
        v = (1 + p[:, 0]) ** frm - 0.5 * p[:, -1] ** frm
        return v + (p[:, 0] ** (frm - 1)) * p[:, -1] if frm >= 2 else v
    xc = np.broadcast_to(poly(cn), (E, Q, nf)).copy()
    yf = T.interpolate(Id, torch.from_numpy(xc.ravel()).to(cuda), Q).cpu().numpy().reshape(E, Q, nt)
    assert np.abs(yf - poly(fn)).max() <= 1e-12
