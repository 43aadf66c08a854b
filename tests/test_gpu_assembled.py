"""Assembled operator (through the C ABI): the block rows probed from the matrix-free apply against
the oracle's dense assembly, and the assembled apply against the matrix-free apply.

The reference fills a PetscDGMatrix (BAIJ: per element its own block and one block per face
neighbour, app/common/PetscDGMatrix.cpp:14-77) through DGOperator::assemble (src/form/DGOperator.h:
121-170) with assemble_volume / skeleton / boundary (app/localoperator/Elasticity.cpp:317-519,
Poisson.cpp:225-381).  The oracle restates that assembly independently of its matrix-free apply
(the two are pinned against each other on the CPU), so every block of every element must match it
to rounding: a probe that mixed two elements of one colour, dropped a face or put a block in the
wrong slot fails here.  Bars: blocks 1e-12 relative to the largest entry; the FP64 assembled apply
equals the matrix-free FP64 apply to 1e-13 (relative L2), the FP32 one to 2e-6; every column of a
batched apply equals its single-column apply exactly and the padding between columns is untouched.
"""
import ctypes as C

import numpy as np
import pytest

from cases import Case, rel_l2
from tandem_amd._lib import lib

pytestmark = pytest.mark.gpu

BC_D, BC_N = 5, 1
# (dim, cells, scheme, degree, kind, bc): the register kernel's block sizes 6, 12 (2D elasticity P1, P2),
# 12 with five slots (3D elasticity P1), 3, 4, 6 (Poisson), and the generic kernel at 20 and 105
CASES = [
    (2, 4, 0, 1, 0, BC_D),
    (2, 3, 0, 2, 0, [BC_N, BC_D, BC_D, BC_N]),
    (3, 2, 1, 1, 0, BC_D),
    (2, 4, 0, 1, 1, BC_D),
    (3, 2, 0, 1, 1, [BC_D, BC_N, BC_D, BC_D, BC_N, BC_D]),
    (2, 3, 0, 2, 1, BC_D),
    (2, 2, 0, 3, 0, BC_D),
    (3, 1, 1, 4, 0, BC_D),
]
IDS = ["D%d-n%d-s%d-P%d-%s" % (c[0], c[1], c[2], c[3], "elast" if c[4] == 0 else "poisson") for c in CASES]


def neighbours(op):
    """nb[e][f]: the element across local face f of e (e itself on the boundary), from the facet table."""
    fac = op.facets()
    nb = np.full((op.E, op.dim + 1), -1, dtype=np.int64)
    for e0, e1, l0, l1, _ in fac:
        nb[e0, l0] = e1
        if e1 != e0:
            nb[e1, l1] = e0
    assert np.all(nb >= 0)
    return nb


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_assembled_blocks_and_apply(cuda, case):
    import torch

    D, n, scheme, N, kind, bc = case
    c = Case(D, n, scheme, N, kind, bc=bc, seed=9)
    op = c.op
    colours = op.assemble()
    A = c.oracle.assemble()
    bs, E = op.bs, op.E
    nb = neighbours(op)
    print("case", case, "elements", E, "bs", bs, "colours", colours, flush=True)
    assert colours <= 1 + (D + 1) + (D + 1) * D  # distance-2 greedy bound
    scale = np.abs(A).max()
    worst = 0.0
    for e in range(E):
        rows = slice(e * bs, (e + 1) * bs)
        got = op.assembled_block(e, 0)
        worst = max(worst, np.abs(got - A[rows, rows]).max())
        for f in range(D + 1):
            got = op.assembled_block(e, 1 + f)
            o = nb[e, f]
            ref = np.zeros((bs, bs)) if o == e else A[rows, o * bs:(o + 1) * bs]
            worst = max(worst, np.abs(got - ref).max())
    print("  blocks: max |A_asm - A_oracle| / max |A| = %.3e" % (worst / scale), flush=True)
    assert worst <= 1e-12 * scale

    rng = np.random.default_rng(4)
    x = torch.from_numpy(rng.uniform(-1, 1, c.n)).to(cuda)
    y_mf = torch.zeros_like(x)
    op.apply(x, y_mf)
    op.set_variant("assembled")
    y = torch.zeros_like(x)
    op.apply(x, y)
    torch.cuda.synchronize()
    e64 = rel_l2(y.cpu().numpy(), y_mf.cpu().numpy())
    e_or = rel_l2(y.cpu().numpy(), A @ x.cpu().numpy())
    x32 = x.float()
    y32 = torch.zeros_like(x32)
    op.apply(x32, y32)
    torch.cuda.synchronize()
    e32 = rel_l2(y32.double().cpu().numpy(), y_mf.cpu().numpy())
    print("  apply: FP64 vs matrix-free %.3e, vs oracle A x %.3e, FP32 %.3e" % (e64, e_or, e32), flush=True)
    assert e64 <= 1e-13 and e_or <= 1e-12 and e32 <= 2e-6

    # ragged column count at a padded stride: each column is its single-column apply, the padding stays
    nc, ld = 19, c.n + 5
    X = torch.from_numpy(rng.uniform(-1, 1, (nc, ld))).to(cuda)
    Y = torch.full((nc, ld), 7.0, dtype=torch.float64, device=cuda)
    op.apply_batched(X, Y)
    for col in (0, 7, 18):
        yc = torch.zeros(c.n, dtype=torch.float64, device=cuda)
        op.apply(X[col, :c.n].contiguous(), yc)
        assert torch.equal(Y[col, :c.n], yc)
    assert torch.all(Y[:, c.n:] == 7.0)

    # the allocation-free entry point runs the assembled apply too
    y2 = torch.zeros_like(x)
    assert lib().tdg_mult_noalloc(op._h, C.c_void_p(x.data_ptr()), C.c_void_p(y2.data_ptr())) == 0
    torch.cuda.synchronize()
    assert torch.equal(y2, y)


def test_assembled_many_columns_and_material_change(cuda):
    """64 columns of a P1 level (the register kernel splits them into column ranges) against the
    matrix-free apply, and a new material drops the blocks: the next apply re-assembles."""
    import torch

    c = Case(2, 16, 0, 1, 0, seed=3)
    op = c.op
    n = c.n
    rng = np.random.default_rng(6)
    nc, ld = 64, n + 1
    X = torch.from_numpy(rng.uniform(-1, 1, (nc, ld))).to(cuda)
    ref = torch.zeros_like(X)
    op.apply_batched(X, ref)
    op.set_variant("assembled")
    for dtype, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        Xd = X.to(dtype)
        Y = torch.zeros_like(Xd)
        op.apply_batched(Xd, Y)
        torch.cuda.synchronize()
        err = max(rel_l2(Y[k, :n].double().cpu().numpy(), ref[k, :n].cpu().numpy()) for k in range(nc))
        print(dtype, "64 columns: worst column rel L2 vs matrix-free %.3e" % err, flush=True)
        assert err <= tol
    lam, mu = 2.0 * c.c0, 0.5 * c.c1
    op.set_material(lam, mu)
    y = torch.zeros(n, dtype=torch.float64, device=cuda)
    op.apply(X[0, :n].contiguous(), y)
    op.set_variant("auto")
    y_mf = torch.zeros_like(y)
    op.apply(X[0, :n].contiguous(), y_mf)
    torch.cuda.synchronize()
    err = rel_l2(y.cpu().numpy(), y_mf.cpu().numpy())
    print("after set_material: assembled vs matrix-free %.3e" % err, flush=True)
    assert err <= 1e-13
