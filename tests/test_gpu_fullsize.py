"""The P4 tet elasticity apply at its full benchmark size (24^3 Kuhn cubes x 6 = 82,944 curved
elements, 8,709,120 DoF, warped P4 geometry, seeded log-normal Lame fields), through the C ABI.

Bars: FP64 relative L2 <= 1e-12 against the CPU oracle's apply of the same vector; FP32 <= 1e-5;
the size-independent properties of the SIPG operator: symmetry y.(A x) = x.(A y) and linearity to
1e-12 relative, the rigid-body null space with natural boundaries on every face (natural faces are
skipped, so A r = 0 for the six rigid motions, exactly representable on the P4 geometry), and a
batched apply equal to the column-by-column one, bit for bit; the matrix-core adjoint gradient equal
to the thread-per-point one (FP64 1e-12; FP32 within 1e-4).
"""
import numpy as np
import pytest

import oracle_lib as O
import tandem_amd as T
from tandem_amd import scenario as S

pytestmark = pytest.mark.gpu

CELLS, N, SEED = 24, 4, 20261203


def _operator(bc):
    mesh = T.Mesh.box(3, CELLS, scheme=1, bc=bc)
    nodes = S.warp(mesh.affine_nodes(N))
    op = T.DGOperator(mesh, T.ELASTICITY, N, nodes)
    lam, mu = S.lame_fields(3, SEED)
    X = op.volume_points()
    op.set_material(lam(X), mu(X))
    return op, nodes


@pytest.fixture(scope="module")
def box24(cuda):
    op, nodes = _operator(T.BC_DIRICHLET)
    print("benchmark operator: %d elements, %d DoF" % (op.E, op.n), flush=True)
    return op, nodes


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_full_size_apply_matches_oracle(cuda, box24):
    import torch

    op, nodes = box24
    om = O.Mesh.box(3, np.broadcast_to(CELLS, (3,)), 1, [T.BC_DIRICHLET] * 6)
    vc, _ = om.coords(N, nodes)
    lam, mu = S.lame_fields(3, SEED)
    oo = O.Operator(0, om, N, lam(vc), mu(vc), nodes=nodes)
    assert oo.n == op.n
    x = np.random.default_rng(SEED).uniform(-1.0, 1.0, op.n)
    ref = oo.apply(x)
    print("oracle apply done", flush=True)
    xd = torch.from_numpy(x).to(cuda)
    y = torch.full_like(xd, float("nan"))
    op.apply(xd, y)
    e64 = _rel(y.cpu().numpy(), ref)
    yf = torch.full((op.n,), float("nan"), dtype=torch.float32, device=cuda)
    op.apply(xd.float(), yf)
    e32 = _rel(yf.double().cpu().numpy(), ref)
    print("full-size rel L2 vs oracle: fp64 %.3e  fp32 %.3e" % (e64, e32), flush=True)
    assert e64 <= 1e-12
    assert e32 <= 1e-5
    # every launch shape and workgroup order of the two-kernel apply; at this size the persistent
    # order (2) runs fewer workgroups than blocks, so each workgroup loops over several blocks
    saved = {p: op.autotune_shapes(1, p)[0] for p in ("fp64", "fp32")}
    for prec, dt, bar in (("fp64", torch.float64, 1e-12), ("fp32", torch.float32, 1e-5)):
        xs = xd.to(dt)
        for shape in range(T.operator.SPLIT_SHAPES):
            for order in range(T.operator.SPLIT_ORDERS):
                op.set_shape(prec, shape, order)
                yy = torch.full_like(xs, float("nan"))
                op.apply(xs, yy)
                err = _rel(yy.double().cpu().numpy(), ref)
                print("  %s shape %d order %d: rel L2 %.3e" % (prec, shape, order, err), flush=True)
                assert err <= bar
        op.set_shape(prec, *saved[prec])


def test_full_size_symmetry_linearity_and_batching(cuda, box24):
    import torch

    op, _ = box24
    rng = np.random.default_rng(SEED + 1)
    X = torch.from_numpy(rng.uniform(-1.0, 1.0, (3, op.n))).to(cuda)
    Y = torch.empty_like(X)
    for i in range(3):
        op.apply(X[i], Y[i])
    x, y = X[0], X[1]
    Ax, Ay = Y[0], Y[1]
    s1, s2 = float(torch.dot(y, Ax)), float(torch.dot(x, Ay))
    scale = float(torch.linalg.norm(y) * torch.linalg.norm(Ax))
    print("symmetry |y.Ax - x.Ay| / scale = %.3e" % (abs(s1 - s2) / scale), flush=True)
    assert abs(s1 - s2) <= 1e-12 * scale
    z = 2.0 * x - 3.0 * y
    Az = torch.empty_like(z)
    op.apply(z, Az)
    lin = float(torch.linalg.norm(Az - (2.0 * Ax - 3.0 * Ay)) / torch.linalg.norm(Az))
    print("linearity %.3e" % lin, flush=True)
    assert lin <= 1e-12
    YB = torch.full_like(X, float("nan"))
    op.apply_batched(X, YB)
    assert torch.equal(YB, Y)


def test_full_size_rigid_body_null_space(cuda):
    import torch

    op, nodes = _operator(T.BC_NATURAL)
    rng = np.random.default_rng(SEED + 2)
    xr = torch.from_numpy(rng.uniform(-1.0, 1.0, op.n)).to(cuda)
    yr = torch.empty_like(xr)
    op.apply(xr, yr)
    a_scale = float(torch.linalg.norm(yr) / torch.linalg.norm(xr))  # |A x| / |x| of a random x
    modes = [np.broadcast_to(e, nodes.shape) for e in np.eye(3)] + [np.cross(w, nodes) for w in np.eye(3)]
    for i, r in enumerate(modes):
        v = torch.from_numpy(np.ascontiguousarray(np.transpose(r, (0, 2, 1))).ravel()).to(cuda)
        y = torch.empty_like(v)
        op.apply(v, y)
        ratio = float(torch.linalg.norm(y) / (a_scale * torch.linalg.norm(v)))
        print("rigid mode %d: |A r| / (|A| |r|) = %.3e" % (i, ratio), flush=True)
        assert ratio <= 1e-11


def test_full_size_gradient_kernels_agree(cuda, box24):
    """The matrix-core adjoint gradient equals the thread-per-point kernel (an independent
    implementation, checked against the oracle on small meshes) at full size: FP64 to 1e-12
    relative, and the FP32 matrix-core gradient stays within 1e-4 of it."""
    import torch

    op, _ = box24
    rng = np.random.default_rng(SEED + 3)
    u = torch.from_numpy(rng.uniform(-1.0, 1.0, op.n)).to(cuda)
    psi = torch.from_numpy(rng.uniform(-1.0, 1.0, op.n)).to(cuda)
    try:
        op.set_variant("points")
        ref = torch.cat([g.clone() for g in op.material_gradient(u, psi)])
        op.set_variant("auto")
        g64 = torch.cat([g.clone() for g in op.material_gradient(u, psi)])
        g32 = torch.cat([g.double() for g in op.material_gradient(u.float(), psi.float())])
    finally:
        op.set_variant("auto")
    e64 = float(torch.linalg.norm(g64 - ref) / torch.linalg.norm(ref))
    e32 = float(torch.linalg.norm(g32 - ref) / torch.linalg.norm(ref))
    print("gradient at full size: mfma fp64 vs points %.3e, mfma fp32 %.3e" % (e64, e32), flush=True)
    assert e64 <= 1e-12
    assert e32 <= 1e-4
