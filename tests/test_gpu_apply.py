"""Parity of the HIP operator apply (through the C ABI) with the CPU FP64 oracle.

Bar: relative L2 difference <= 1e-12 in FP64 (the parity bar of the plain apply);
<= 1e-5 for the FP32 kernels (FP32 copies of the FP64 precompute, FP32 arithmetic).
"""
import numpy as np
import pytest

from cases import BC_D, BC_N, Case, rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # (dim, cells, scheme, degree, kind, bc)
    (3, 2, 1, 4, 0, BC_D),          # P4 tet elasticity, Kuhn cubes (the benchmark element type)
    (3, 2, 0, 4, 0, [BC_D, BC_N, BC_N, BC_D, BC_D, BC_N]),  # tandem GenMesh 5-tet cubes, mixed BCs
    (3, 2, 1, 1, 0, BC_D),
    (3, 2, 1, 2, 0, BC_D),
    (3, 2, 1, 3, 0, BC_D),
    (2, 4, 0, 2, 0, BC_D),
    (2, 4, 0, 3, 0, [BC_N, BC_D, BC_D, BC_D]),  # P3 triangles (the middle p-multigrid level)
    (2, 3, 0, 6, 0, [BC_D, BC_N, BC_D, BC_N]),  # P6 triangles (the Green's-function element type)
    (2, 32, 0, 2, 1, BC_D),         # Poisson P2 on the 32x32x2 unit square
    (2, 3, 0, 4, 1, [BC_D, BC_N, BC_N, BC_D]),
    (3, 2, 1, 2, 1, BC_D),
    (3, 2, 0, 4, 1, BC_D),
]


def _ids(c):
    return "D%d-n%s-s%d-P%d-%s" % (c[0], c[1], c[2], c[3], "elast" if c[4] == 0 else "poisson")


@pytest.mark.parametrize("variant", ["auto", "points", "mfma-xcd", "mfma-element"])
@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_apply_matches_oracle(cuda, case, variant):
    """variant "auto" runs the two-kernel matrix-core apply (facet kernel + element kernel) where
    one is compiled (elasticity 3D P2-P4, 2D P2-P4 and P6), "points" the thread-per-point kernel,
    "mfma-element" the element-centric matrix-core kernel and "mfma-xcd" the same with its
    workgroups renumbered per XCD."""
    import torch

    D, n, scheme, N, kind, bc = case
    c = Case(D, n, scheme, N, kind, bc=bc, seed=7)
    c.op.set_variant(variant)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1.0, 1.0, c.n)
    y_ref = c.oracle.apply(x, nthreads=4)
    print("case", _ids(case), variant, "n", c.n, flush=True)
    xd = torch.from_numpy(x).to(cuda)
    yd = torch.full_like(xd, float("nan"))
    c.op.apply(xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    assert np.all(np.isfinite(y))
    err64 = rel_l2(y, y_ref)
    xf = xd.float()
    yf = torch.full_like(xf, float("nan"))
    c.op.apply(xf, yf)
    torch.cuda.synchronize()
    err32 = rel_l2(yf.double().cpu().numpy(), y_ref)
    print("  fp64 rel L2 %.3e  fp32 rel L2 %.3e" % (err64, err32), flush=True)
    assert err64 <= 1e-12
    assert err32 <= 1e-5


@pytest.mark.parametrize("case", [(3, 2, 1, 4), (3, 3, 1, 2), (2, 5, 0, 6)], ids=["D3-P4", "D3-P2", "D2-P6"])
def test_batched_apply_equals_columns(cuda, case):
    """Batched columns (grid.y) and a ragged last 16-element group: each column of the batched apply
    equals the single apply bit for bit."""
    import torch

    D, n, scheme, N = case
    c = Case(D, n, scheme, N, 0, seed=3)
    rng = np.random.default_rng(5)
    ncols, ld = 5, c.n + 7
    X = torch.from_numpy(rng.uniform(-1, 1, (ncols, ld))).to(cuda)
    Y = torch.zeros_like(X)
    c.op.apply_batched(X, Y)
    for j in range(ncols):
        y = torch.zeros(c.n, dtype=torch.float64, device=cuda)
        c.op.apply(X[j, :c.n].contiguous(), y)
        assert torch.equal(y, Y[j, :c.n])
    torch.cuda.synchronize()
    Xf, Yf = X.float(), torch.zeros(ncols, ld, dtype=torch.float32, device=cuda)
    c.op.apply_batched(Xf, Yf)
    y = torch.zeros(c.n, dtype=torch.float32, device=cuda)
    c.op.apply(Xf[2, :c.n].contiguous(), y)
    assert torch.equal(y, Yf[2, :c.n])


def test_apply_is_deterministic(cuda):
    import torch

    c = Case(3, 2, 1, 4, 0, seed=4)
    x = torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, c.n)).to(cuda)
    y1, y2 = torch.zeros_like(x), torch.zeros_like(x)
    c.op.apply(x, y1)
    c.op.apply(x, y2)
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotuned_variant_matches_oracle(cuda, precision):
    """The autotuner times the applicable variants, keeps one, and the apply it leaves in place
    still meets the parity bar."""
    import torch

    c = Case(3, 2, 1, 4, 0, seed=6)
    name, ms = c.op.autotune(ncols=2, precision=precision)
    print("autotune", precision, name, ms, flush=True)
    assert name in ms and set(ms) == {"auto", "mfma-element", "mfma-xcd", "points"}
    assert all(v > 0 for v in ms.values())
    x = np.random.default_rng(7).uniform(-1, 1, c.n)
    ref = c.oracle.apply(x)
    dt = torch.float64 if precision == "fp64" else torch.float32
    xd = torch.from_numpy(x).to(cuda).to(dt)
    y = torch.zeros_like(xd)
    c.op.apply(xd, y)
    assert rel_l2(y.double().cpu().numpy(), ref) <= (1e-12 if precision == "fp64" else 1e-5)
