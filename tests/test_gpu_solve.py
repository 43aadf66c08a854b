"""GPU Krylov solves through the C ABI: FP64 CG and FP32 CG inside FP64 iterative refinement.

Checks: both reach the FP64 true relative residual 1e-10; their solutions agree; the FP64 solution
matches a dense solve of the oracle's assembled matrix (independent restatement of the bilinear
form) on a small mesh.  (The convergence orders of the manufactured cosine solution are checked in
test_gpu_convergence.py.)
"""
import numpy as np
import pytest

from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


def _solve(op, b, mode, **kw):
    import torch

    bd = torch.from_numpy(b).cuda()
    x = torch.zeros_like(bd)
    r = op.solve(bd, x, mode=mode, **kw)
    return x.cpu().numpy(), r


def test_cg_matches_dense_solve(cuda):
    c = Case(3, 2, 1, 2, 0, seed=2)
    b = np.random.default_rng(3).uniform(-1, 1, c.n)
    A = c.oracle.assemble()
    x_ref = np.linalg.solve(A, b)
    for mode in ("fp64", "fp32-ir", "fp32-ru"):
        x, r = _solve(c.op, b, mode, rtol=1e-10)
        print(mode, r, flush=True)
        assert r["converged"] and r["rel_residual"] <= 1e-10
        assert rel_l2(x, x_ref) <= 1e-7


def test_poisson_unit_square_solve(cuda):
    """Poisson P2 on the 32x32x2 unit square, u = sin(pi x) sin(pi y), CG to 1e-10."""
    import tandem_amd as T
    from tandem_amd import scenario as S

    mesh = T.Mesh.box(2, 32, 0, T.BC_DIRICHLET)
    op = T.Poisson(mesh, 2, K=1.0)
    X = op.volume_points()
    _, f = S.poisson_sine(X)
    g = np.zeros((op.F, 1, op.nq))
    b = op.rhs(f[:, None, :], g)
    x64, r64 = _solve(op, b, "fp64", rtol=1e-10)
    for mode in ("fp32-ir", "fp32-ru"):
        x32, r32 = _solve(op, b, mode, rtol=1e-10)
        print("fp64", r64, mode, r32, flush=True)
        assert r64["converged"] and r32["converged"]
        assert rel_l2(x32, x64) <= 1e-8


@pytest.mark.parametrize("mode,precond", [("fp64", None), ("fp32-ru", None), ("fp32-ru", "block-jacobi"),
                                         ("fp64", "block-jacobi-f32")])
def test_multi_column_solve_matches_dense(cuda, mode, precond):
    """ncols right-hand sides in lockstep (one batched apply per iteration): every column reaches the
    FP64 true relative residual 1e-10 and the dense solution; a zero column stays zero."""
    import torch

    c = Case(3, 2, 1, 2, 0, seed=4)
    A = c.oracle.assemble()
    nc, ld = 5, c.n + 3
    Bh = np.random.default_rng(7).uniform(-1, 1, (nc, ld))
    Bh[3] = 0.0
    Bh[1] *= 1e3  # columns of different scale converge independently
    B = torch.from_numpy(Bh).to(cuda)
    X = torch.zeros_like(B)
    r = c.op.solve_multi(B, X, rtol=1e-10, mode=mode, precond=precond)
    print(mode, precond, r, flush=True)
    assert r["converged"]
    Xh = X.cpu().numpy()
    for j in range(nc):
        if j == 3:
            assert np.all(Xh[j] == 0.0)
            continue
        assert r["column_rel_residual"][j] <= 1e-10
        assert rel_l2(Xh[j, :c.n], np.linalg.solve(A, Bh[j, :c.n])) <= 1e-7


@pytest.mark.parametrize("variant", ["auto", "assembled"])
def test_multi_column_solve_small_blocks(cuda, variant):
    """P1 operators (the p-multigrid coarse level): the block-Jacobi application with its fused r . z
    partial sums (small-block kernel) inside lockstep CG, on the matrix-free and the assembled apply,
    FP64 and FP32 with FP64 replacement, 18 columns (a ragged second range of the 16-column kernel),
    replayed from graphs and launched: every column reaches 1e-10 and the dense solution."""
    import torch

    c = Case(2, 6, 0, 1, 0, seed=12)
    c.op.set_variant(variant)
    A = c.oracle.assemble()
    nc, ld = 18, c.n + 2
    Bh = np.random.default_rng(13).uniform(-1, 1, (nc, ld))
    B = torch.from_numpy(Bh).to(cuda)
    for mode in ("fp64", "fp32-ru"):
        its = {}
        for graphs in (False, True):
            X = torch.zeros_like(B)
            r = c.op.solve_multi(B, X, rtol=1e-10, mode=mode, precond="block-jacobi", graphs=graphs)
            print(variant, mode, graphs, r["iterations"], max(r["column_rel_residual"]), flush=True)
            assert r["converged"]
            Xh = X.cpu().numpy()
            for j in (0, 9, 17):
                assert rel_l2(Xh[j, :c.n], np.linalg.solve(A, Bh[j, :c.n])) <= 1e-7
            its[graphs] = (r["iterations"], Xh)
        assert its[True][0] == its[False][0] and np.array_equal(its[True][1], its[False][1])


@pytest.mark.parametrize("mode,precond", [("fp64", None), ("fp32-ru", "block-jacobi"), ("fp32-ir", None)])
def test_graph_replayed_solve_equals_launched(cuda, mode, precond):
    """The iterations between residual checks replayed from a captured HIP graph (operator on a
    non-default stream) give bit-identical iterates to plain launches, over several replays."""
    import torch

    c = Case(3, 2, 1, 3, 0, seed=5)
    stream = torch.cuda.Stream()
    c.op.set_stream(stream)
    b = torch.from_numpy(np.random.default_rng(6).uniform(-1, 1, c.n)).to(cuda)
    out = {}
    for graphs in (False, True):
        x = torch.zeros_like(b)
        torch.cuda.synchronize()  # b and x are made on the default stream; the solve runs on `stream`
        with torch.cuda.stream(stream):
            r = c.op.solve(b, x, rtol=1e-10, mode=mode, precond=precond, graphs=graphs)
        stream.synchronize()
        print(mode, precond, graphs, r, flush=True)
        assert r["converged"]
        out[graphs] = (x.cpu().numpy(), r["iterations"])
    assert out[True][1] == out[False][1]
    assert np.array_equal(out[True][0], out[False][0])
    c.op.set_stream(None)


@pytest.mark.parametrize("mode,precond", [("fp64", None), ("fp32-ru", "block-jacobi")])
def test_graph_replayed_multi_solve_equals_launched(cuda, mode, precond):
    """Lockstep multi-column and segmented solves with their iteration chunks replayed from
    captured HIP graphs give bit-identical iterates to plain launches, over several replays."""
    import torch
    import tandem_amd as T

    c = Case(2, 3, 0, 4, 0, seed=7)
    stream = torch.cuda.Stream()
    c.op.set_stream(stream)
    rng = np.random.default_rng(8)
    B = torch.from_numpy(rng.uniform(-1, 1, (3, c.n + 5))).to(cuda)
    out = {}
    for graphs in (False, True):
        X = torch.zeros_like(B)
        torch.cuda.synchronize()  # B and X are made on the default stream; the solve runs on `stream`
        with torch.cuda.stream(stream):
            r = c.op.solve_multi(B, X, rtol=1e-10, mode=mode, precond=precond, graphs=graphs)
        stream.synchronize()
        print("multi", mode, precond, graphs, r["iterations"], r["column_iterations"], flush=True)
        assert r["converged"]
        out[graphs] = (X.cpu().numpy(), r["iterations"])
    assert out[True][1] == out[False][1]
    assert np.array_equal(out[True][0], out[False][0])
    c.op.set_stream(None)
    # segmented: three boxes in one operator
    meshes = []
    for i, n in enumerate((2, 4, 3)):
        m = T.Mesh.box(2, n, 0, T.BC_DIRICHLET)
        meshes.append(T.Mesh(m.vertices + np.array([2.0 * i, 0.0]), m.elements, m.boundary_facets, m.boundary_bc))
    union, eoff = T.Mesh.union(meshes)
    op = T.DGOperator(union, T.ELASTICITY, 4)
    op.set_material(np.full((op.E, op.Nq), 2.0), np.full((op.E, op.Nq), 1.0))
    op.set_stream(stream)
    b = torch.from_numpy(rng.uniform(-1, 1, op.n)).to(cuda)
    seg = {}
    for graphs in (False, True):
        x = torch.zeros_like(b)
        torch.cuda.synchronize()  # b and x are made on the default stream; the solve runs on `stream`
        with torch.cuda.stream(stream):
            r = op.solve_segments(b, x, eoff * op.bs, rtol=1e-10, mode=mode, precond=precond, graphs=graphs)
        stream.synchronize()
        print("segments", mode, precond, graphs, r["iterations"], r["segment_iterations"], flush=True)
        assert r["converged"]
        seg[graphs] = (x.cpu().numpy(), r["iterations"])
    assert seg[True][1] == seg[False][1]
    assert np.array_equal(seg[True][0], seg[False][0])
    op.set_stream(None)
