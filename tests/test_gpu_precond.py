"""Element-block Jacobi preconditioner (through the C ABI) against the oracle's assembled matrix,
and preconditioned solves.

The oracle's dense assembly (assemble_volume/skeleton/boundary restated, pinned against the
matrix-free apply) gives the exact diagonal blocks: the GPU inverse blocks must match inv(A_KK) to
1e-10 relative, the block application D^-1 r likewise, and preconditioned CG (FP64 and FP32 with
FP64 residual replacement) must reach the FP64 true relative residual 1e-10 in fewer iterations
than plain CG, with the dense solution.
"""
import numpy as np
import pytest

from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

# block sizes 105, 30, 56 (one workgroup per element) and 6, 10, 6, 12 (the small-block kernel)
CASES = [(3, 2, 1, 4, 0), (3, 2, 0, 2, 0), (2, 3, 0, 6, 0), (2, 4, 0, 2, 1), (3, 2, 1, 2, 1), (2, 4, 0, 1, 0),
         (3, 2, 1, 1, 0)]
IDS = ["D%d-n%d-s%d-P%d-%s" % (c[:4] + ("elast" if c[4] == 0 else "poisson",)) for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_block_inverses_match_assembled_diagonal(cuda, case):
    import torch

    D, n, scheme, N, kind = case
    c = Case(D, n, scheme, N, kind, seed=6)
    ncol = c.op.block_jacobi_build()
    A = c.oracle.assemble()
    bs, E = c.op.bs, c.op.E
    print("case", case, "colours", ncol, "bs", bs, flush=True)
    assert 2 <= ncol <= 8
    worst = 0.0
    for e in sorted({0, E // 3, E // 2, E - 1}):
        blk = A[e * bs:(e + 1) * bs, e * bs:(e + 1) * bs]
        ref = np.linalg.inv(blk)
        got = c.op.block_jacobi_inverse(e)
        worst = max(worst, np.abs(got - ref).max() / np.abs(ref).max())
    print("  inverse block max rel %.3e" % worst, flush=True)
    assert worst <= 1e-10
    r = np.random.default_rng(2).uniform(-1, 1, c.n)
    ref = np.concatenate([np.linalg.solve(A[e * bs:(e + 1) * bs, e * bs:(e + 1) * bs], r[e * bs:(e + 1) * bs])
                          for e in range(E)])
    z = torch.zeros(c.n, dtype=torch.float64, device=cuda)
    c.op.block_jacobi_apply(torch.from_numpy(r).to(cuda), z)
    err = rel_l2(z.cpu().numpy(), ref)
    print("  D^-1 r rel L2 %.3e" % err, flush=True)
    assert err <= 1e-10
    # several columns at a stride: each equals the single-column application (19 columns: a ragged
    # last pass of the 16-column staging and of the 4-column lane tiles)
    nc, ld = 19, c.n + 3
    R = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, (nc, ld))).to(cuda)
    Z = torch.full((nc, ld), 7.0, dtype=torch.float64, device=cuda)
    c.op.block_jacobi_apply_batched(R, Z)
    for col in (0, 5, 15, 16, 18):
        zc = torch.zeros(c.n, dtype=torch.float64, device=cuda)
        c.op.block_jacobi_apply(R[col, :c.n].contiguous(), zc)
        assert rel_l2(Z[col, :c.n].cpu().numpy(), zc.cpu().numpy()) <= 1e-14
    assert torch.all(Z[:, c.n:] == 7.0)


def test_preconditioned_solves(cuda):
    import torch

    c = Case(3, 3, 1, 3, 0, seed=2)
    b = np.random.default_rng(3).uniform(-1, 1, c.n)
    x_ref = np.linalg.solve(c.oracle.assemble(), b)
    bd = torch.from_numpy(b).to(cuda)
    its = {}
    for mode in ("fp64", "fp32-ru", "fp32-ir"):
        for pc in (None, "block-jacobi", "block-jacobi-f32"):
            x = torch.zeros_like(bd)
            r = c.op.solve(bd, x, rtol=1e-10, mode=mode, precond=pc)
            print(mode, pc, r, flush=True)
            assert r["converged"] and r["rel_residual"] <= 1e-10
            assert rel_l2(x.cpu().numpy(), x_ref) <= 1e-7
            its[(mode, pc)] = r["iterations"]
    assert its[("fp64", "block-jacobi")] < its[("fp64", None)]
    # the FP32-stored blocks are the same preconditioner to FP32 rounding: the FP64 solve's iteration
    # count barely moves, and FP32 solves (which always apply them) are unchanged
    assert abs(its[("fp64", "block-jacobi-f32")] - its[("fp64", "block-jacobi")]) <= 0.05 * its[("fp64", "block-jacobi")]
    assert its[("fp32-ru", "block-jacobi-f32")] == its[("fp32-ru", "block-jacobi")]
    assert its[("fp32-ru", "block-jacobi")] < its[("fp32-ru", None)]
