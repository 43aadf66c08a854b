"""Host side of the MI355X operator (no GPU needed): the C ABI library loads and exports every
entry point declared in include/tandem_dg.h; its mesh builder, topology, FP64 geometry and
material precompute and right-hand side agree with the CPU oracle; error handling at the boundary."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import tandem_amd as T
import tandem_amd._lib as L
from cases import warp_nodes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = L.lib()
    hdr = open(os.path.join(ROOT, "include", "tandem_dg.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|char const\*)\s+(tdg_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 20
    for n in names:
        assert getattr(lib, n, None) is not None, n
    assert set(L.EXPORTS) <= set(names)


CASES = [(3, 2, 1, 4), (3, 2, 0, 2), (2, 4, 0, 2), (2, 3, 0, 6)]


@pytest.mark.parametrize("D,n,scheme,N", CASES)
def test_mesh_and_precompute_match_oracle(D, n, scheme, N):
    bcs = [5] * (2 * D)
    om = O.Mesh.box(D, [n] * D, scheme, bcs)
    pm = T.Mesh.box(D, n, scheme, 5)
    assert np.array_equal(om.verts, pm.vertices) and np.array_equal(om.elems, pm.elements)
    assert np.array_equal(om.bfacets, pm.boundary_facets) and np.array_equal(om.bbc, pm.boundary_bc)
    assert np.allclose(pm.affine_nodes(N), om.affine_nodes(N), atol=1e-15)
    nodes = warp_nodes(pm.affine_nodes(N), 0.02)
    for kind in (0, 1):
        op = T.DGOperator(pm, kind, N, nodes)
        vc = op.volume_points()
        ovc, ofc = om.coords(N, nodes)
        assert np.array_equal(vc, ovc) and np.array_equal(op.facet_points(), ofc)
        c0 = 2 + 0.3 * np.sin(3 * vc[..., 0]) * np.cos(2 * vc[..., 1])
        c1 = 1 + 0.2 * np.cos(2 * vc[..., D - 1])
        op.set_material(c0, c1 if kind == 0 else None)
        oo = O.Operator(kind, om, N, c0, c1 if kind == 0 else None, nodes=nodes)
        assert np.array_equal(op.facets(), oo.facets())
        for a in ("absDetJ", "JInv", "volume", "normal", "area", "normalLength", "JInv0", "JInv1"):
            ref = oo.array(a)
            assert np.abs(op.host_array(a) - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), a
        pairs = [("penalty", "penalty")] + ([("mat0", "lam"), ("mat1", "mu"), ("c0Q", "lamWJ"), ("c1Q", "muWJ")]
                                           if kind == 0 else [("mat0", "K"), ("c0Q", "JWK")])
        for a, b in pairs:
            ref = oo.array(b)
            assert np.abs(op.host_array(a) - ref).max() <= 1e-11 * np.abs(ref).max(), a
        Q = D if kind == 0 else 1
        f = np.ascontiguousarray(np.sin(vc[..., None, 0] + np.arange(Q)[None, None, :]).transpose(0, 2, 1))
        fc = op.facet_points()
        g = np.ascontiguousarray(np.cos(2 * fc[..., None, 0] + np.arange(Q)).transpose(0, 2, 1))
        rb, ro = op.rhs(f, g), oo.rhs(f, g)
        assert np.abs(rb - ro).max() <= 1e-12 * np.abs(ro).max()


def test_size_queries_agree():
    """tdg_block_size / tdg_local_size (the MatShell's block size and local rows) against tdg_sizes."""
    from tandem_amd._lib import lib

    for D, N, kind, bs in [(3, 4, 0, 105), (2, 6, 0, 56), (2, 2, 1, 6)]:
        op = T.DGOperator(T.Mesh.box(D, 2, 0, 5), kind, N)
        b, n = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
        assert lib().tdg_block_size(op._h, b) == 0 and lib().tdg_local_size(op._h, n) == 0
        assert b[0] == op.bs == bs and n[0] == op.n == bs * op.E
    assert lib().tdg_block_size(None, b) != 0


def test_apply_counts_match_closed_form():
    """Algorithmic bytes / flops of the P4 tet apply per element (the roofline's numerator)."""
    m = T.Mesh.box(3, 4, 1, 5)
    op = T.DGOperator(m, 0, 4)
    B, F = op.apply_counts(8)
    E, Fu = op.E, op.F
    assert F == pytest.approx(E * 79857 + 4 * E * 33972 + Fu * 72 * 19)
    assert B == pytest.approx(8 * (E * (2 * 105 + 57 * 11) + Fu * (18 * 19 + 6 * 19 + 4 * 19 + 1)))


def test_boundary_errors():
    lib = L.lib()
    m = T.Mesh.box(2, 2, 0, 5)
    op = T.DGOperator(m, 0, 2)
    # apply before set_material is a state error, reported as a return code with a message
    rc = lib.tdg_mult(op._h, C.c_void_p(16), C.c_void_p(32))
    assert rc == -2 and b"set_material" in lib.tdg_last_error()
    # null arguments
    assert lib.tdg_mult(None, None, None) == -1
    # unknown kind / bad degree
    with pytest.raises(L.TdgError):
        T.DGOperator(m, 7, 2)
    with pytest.raises(L.TdgError):
        T.DGOperator(m, 0, 9)
    # out-of-range vertex ids
    bad = T.Mesh(m.vertices, np.array([[0, 1, 999]]), None, None)
    with pytest.raises(L.TdgError):
        T.DGOperator(bad, 0, 2)
    # unknown boundary condition value
    bb = T.Mesh(m.vertices, m.elements, m.boundary_facets, np.full(len(m.boundary_bc), 4, dtype=np.int32))
    with pytest.raises(L.TdgError):
        T.DGOperator(bb, 0, 2)


@pytest.mark.parametrize("kind,D,N,scheme", [(0, 3, 2, 1), (0, 2, 4, 0), (1, 2, 2, 0), (1, 3, 2, 0)])
def test_rhs_with_fault_and_boundary_fault_matches_oracle(kind, D, N, scheme):
    """Host rhs (rhs_volume + rhs_skeleton + rhs_boundary) on a mesh with an interior fault plane and
    Fault/Dirichlet/natural box faces against the oracle."""
    from cases import fault_plane_mesh
    bcs = [5, 3, 1, 5, 5, 3][:2 * D]
    pm, om = fault_plane_mesh(D, 2 if D == 3 else 4, scheme, bc=bcs)
    nodes = warp_nodes(pm.affine_nodes(N), 0.02)
    op = T.DGOperator(pm, kind, N, nodes)
    vc = op.volume_points()
    c0 = 2 + 0.3 * np.sin(3 * vc[..., 0]) * np.cos(2 * vc[..., 1])
    c1 = 1 + 0.2 * np.cos(2 * vc[..., D - 1])
    op.set_material(c0, c1 if kind == 0 else None)
    oo = O.Operator(kind, om, N, c0, c1 if kind == 0 else None, nodes=nodes)
    assert np.array_equal(op.facets(), oo.facets())
    assert np.sum(oo.facets()[:, 4] == 3) > 0
    Q = D if kind == 0 else 1
    rng = np.random.default_rng(2)
    f = rng.uniform(-1, 1, (op.E, Q, op.Nq))
    g = rng.uniform(-1, 1, (op.F, Q, op.nq))
    rb, ro = op.rhs(f, g), oo.rhs(f, g)
    assert np.abs(rb - ro).max() <= 1e-12 * np.abs(ro).max()


@pytest.mark.parametrize("D,to,frm", [(3, 4, 2), (3, 2, 1), (3, 4, 3), (2, 6, 3), (2, 3, 2), (2, 2, 1)])
def test_interpolation_matrix_matches_oracle(D, to, frm):
    """Host nodal interpolation matrix (NodalInterpolation.cpp:8-38) equals the oracle's, rows sum
    to one (constants are reproduced) and nested nodes give unit rows."""
    I = T.interpolation_matrix(D, to, frm)
    ref = O.nodal_interpolation(D, to, frm)
    assert I.shape == ref.shape
    assert np.abs(I - ref).max() <= 1e-13
    assert np.abs(I.sum(axis=1) - 1).max() <= 1e-13


def test_union_mesh_keeps_each_mesh_facets_in_order():
    """An operator over the disjoint union of meshes numbers elements and facets mesh by mesh, so a
    mesh's facet data and vector segment are contiguous slices."""
    from tandem_amd import scenario as S
    frs = [S.FaultRectangle.with_elements(n, 5 + i) for i, n in enumerate((300, 800, 500))]
    meshes = [fr.mesh() for fr in frs]
    union, eoff = T.Mesh.union(meshes)
    op = T.DGOperator(union, T.ELASTICITY, 2)
    parts = [T.DGOperator(m, T.ELASTICITY, 2).facets() for m in meshes]
    ref = []
    for i, f in enumerate(parts):
        g = f.copy()
        g[:, 0:2] += eoff[i]
        ref.append(g)
    assert np.array_equal(op.facets(), np.concatenate(ref))
    assert list(eoff) == list(np.cumsum([0] + [m.num_elements for m in meshes]))


def test_segmented_solve_rejects_coupled_segments():
    """Segments must be whole elements of meshes with no facet between them (host-side checks run
    before any device work)."""
    m = T.Mesh.box(2, 4, 0, T.BC_DIRICHLET)
    op = T.DGOperator(m, T.ELASTICITY, 2)
    op.set_material(np.full((op.E, op.Nq), 2.0), np.full((op.E, op.Nq), 1.0))
    half = (op.E // 2) * op.bs
    with pytest.raises(L.TdgError, match="couples two segments"):
        L.check(L.lib().tdg_solve_segments(op._h, C.c_void_p(8), C.c_void_p(16), 2,
                                           np.array([0, half, op.n], dtype=np.int64),
                                           C.byref(L.SolveOpts(1e-10, 10, 0, 1e-4, 1, 10, 0.1, 0, 0)),
                                           C.byref(L.SolveResult()), np.zeros(2, dtype=np.int32), np.zeros(2)))
    with pytest.raises(L.TdgError, match="whole elements"):
        L.check(L.lib().tdg_solve_segments(op._h, C.c_void_p(8), C.c_void_p(16), 2,
                                           np.array([0, half + 1, op.n], dtype=np.int64),
                                           C.byref(L.SolveOpts(1e-10, 10, 0, 1e-4, 1, 10, 0.1, 0, 0)),
                                           C.byref(L.SolveResult()), np.zeros(2, dtype=np.int32), np.zeros(2)))


def test_more_boundary_errors():
    """Empty and malformed inputs and calls out of order are reported as error codes with a
    message, before any device work."""
    lib = L.lib()
    m = T.Mesh.box(2, 2, 0, 5)
    # no elements
    empty = T.Mesh(m.vertices, np.zeros((0, 3), dtype=np.int64), None, None)
    with pytest.raises(L.TdgError):
        T.DGOperator(empty, 0, 2)
    # a boundary facet that is not a face of any element
    bad = T.Mesh(m.vertices, m.elements, np.array([[0, 8]]), np.array([5], dtype=np.int32))
    with pytest.raises(L.TdgError):
        T.DGOperator(bad, 0, 2)
    # p-multigrid transfer needs 1 <= from < to
    with pytest.raises(L.TdgError):
        T.interpolation_matrix(2, 2, 2)
    with pytest.raises(L.TdgError):
        T.interpolation_matrix(3, 2, 0)
    # operations that need set_material first
    op = T.DGOperator(m, 0, 2)
    assert lib.tdg_autotune(op._h, 1, 8, None, np.zeros(4)) == -2
    assert lib.tdg_inverse_mass(op._h, C.c_void_p(16), C.c_void_p(32)) == -2
    assert lib.tdg_fault_slip(op._h, C.c_void_p(16), C.c_void_p(32), 1, 0, 0) == -2
    # the fault adapter is built for elasticity
    pois = T.DGOperator(m, 1, 2)
    with pytest.raises(L.TdgError, match="elasticity"):
        pois.fault_setup([0.0, 1.0], [1.0, 0.0])
    # the solver rejects a bad column layout before touching the device
    op.set_material(np.full((op.E, op.Nq), 2.0), np.full((op.E, op.Nq), 1.0))
    opts = L.SolveOpts(1e-10, 10, 0, 1e-4, 1, 10, 0.1, 0, 0)
    res = L.SolveResult()
    rc = lib.tdg_solve_multi(op._h, C.c_void_p(16), C.c_void_p(32), 2, op.n - 1, C.byref(opts), C.byref(res),
                             np.zeros(2, dtype=np.int32), np.zeros(2))
    assert rc == -1 and b"bad" in lib.tdg_last_error()


def test_tensor_arguments_are_checked_before_the_library_is_called():
    """Every entry point of the Python mirror validates dtype, contiguity, size and residency of its
    tensors and raises before handing a pointer to the library (a float32 vector given to an FP64
    entry point would otherwise be read and written as n doubles)."""
    import torch

    m = T.Mesh.box(2, 2, 0, 5)
    op = T.DGOperator(m, 0, 2)
    op.set_material(np.full((op.E, op.Nq), 2.0), np.full((op.E, op.Nq), 1.0))
    n = op.n
    f64, f32 = torch.zeros(n, dtype=torch.float64), torch.zeros(n, dtype=torch.float32)
    strided = torch.zeros(2 * n, dtype=torch.float64)[::2]
    short = torch.zeros(n - 1, dtype=torch.float64)
    i64 = torch.zeros(n, dtype=torch.int64)
    calls = [
        (TypeError, lambda: op.solve(f32, f32)),
        (TypeError, lambda: op.solve(f32, f64)),
        (ValueError, lambda: op.solve(strided, f64)),
        (ValueError, lambda: op.solve(short, f64)),
        (ValueError, lambda: op.solve(f64, f64.clone())),  # host memory
        (TypeError, lambda: op.apply(i64, f64)),
        (ValueError, lambda: op.apply(strided, f64)),
        (ValueError, lambda: op.apply(short, f64)),
        (ValueError, lambda: op.apply_batched(torch.zeros(2, n - 1, dtype=torch.float64), f64)),
        (TypeError, lambda: op.solve_multi(f32.view(1, n), f32.view(1, n))),
        (ValueError, lambda: op.solve_multi(torch.zeros(2, n - 1, dtype=torch.float64),
                                            torch.zeros(2, n - 1, dtype=torch.float64))),
        (ValueError, lambda: op.solve_segments(f64, f64, [0, n + 1])),
        (ValueError, lambda: op.solve_segments(f64, f64, [0, 5, 5, n])),
        (TypeError, lambda: op.solve_segments(f32, f64, [0, n])),
        (ValueError, lambda: op.block_jacobi_apply(strided, f64)),
        (TypeError, lambda: op.block_jacobi_apply(f32, f64)),
        (TypeError, lambda: op.inverse_mass(f32)),
        (ValueError, lambda: op.inverse_mass(short)),
        (TypeError, lambda: op.wave_rhs(f32)),
        (ValueError, lambda: op.traction(short)),
        (TypeError, lambda: op.material_gradient(i64, f64)),
        (TypeError, lambda: op.rhs_device(f32)),
    ]
    for i, (exc, call) in enumerate(calls):
        with pytest.raises(exc):
            call()


def test_pmultigrid_rejects_inconsistent_levels():
    """The p-multigrid hierarchy checks its levels on the host before any device work: at least two,
    one problem on one mesh, strictly decreasing degree, sane options; a replaced penalty needs one
    value per facet."""
    m = T.Mesh.box(2, 3, 0, T.BC_DIRICHLET)
    m2 = T.Mesh.box(2, 4, 0, T.BC_DIRICHLET)

    def make(mesh, p, kind=T.ELASTICITY):
        op = T.DGOperator(mesh, kind, p)
        op.set_material(np.full((op.E, op.Nq), 2.0), None if kind == T.POISSON else np.full((op.E, op.Nq), 1.0))
        return op

    p4, p2, p2b = make(m, 4), make(m, 2), make(m2, 2)
    cases = [
        ([p4], {}, "bad p-multigrid arguments"),
        ([p2, p4], {}, "decreasing degree"),
        ([p4, p4], {}, "decreasing degree"),
        ([p4, p2b], {}, "one mesh"),
        ([p4, make(m, 2, T.POISSON)], {}, "one mesh"),
        ([p4, p2], {"smooth": 0}, "options"),
        ([p4, p2], {"eig_lo": 1.2}, "options"),
        ([p4, p2], {"coarse_rtol": 0.0}, "options"),
    ]
    for levels, kw, msg in cases:
        with pytest.raises(L.TdgError, match=msg):
            T.PMultigrid(levels, **kw)
    unset = T.DGOperator(m, T.ELASTICITY, 2)
    with pytest.raises(L.TdgError, match="set_material"):
        T.PMultigrid([p4, unset])
    # the defaults put the coarsest level on its assembled block rows with the AMG coarse
    # preconditioner (the C struct and its Python mirror agree field by field)
    o = L.PMGOpts()
    L.lib().tdg_pmg_default_opts(C.byref(o))
    assert (o.smooth, o.coarse_max_iter, o.galerkin_penalty, o.coarse_fp32, o.vcycle_fp32) == (6, 500, 1, 1, 1)
    assert (o.eig_lo, o.eig_hi, o.coarse_rtol) == (0.05, 1.1, 0.1)
    assert o.coarse_assembled == 1 and o.coarse_amg == 1


def test_assembled_variant_and_block_arguments_on_the_host():
    """The assembled variant is a valid kernel choice (nothing runs until an apply); block queries
    outside the block row are rejected before any device work."""
    op = T.DGOperator(T.Mesh.box(2, 2, 0, T.BC_DIRICHLET), T.ELASTICITY, 1)
    op.set_material(np.full((op.E, op.Nq), 2.0), np.full((op.E, op.Nq), 1.0))
    op.set_variant("assembled")
    assert L.lib().tdg_set_variant(op._h, 5) != 0
    for e, slot in ((-1, 0), (op.E, 0), (0, -1), (0, op.dim + 2)):
        with pytest.raises(L.TdgError, match="no such block"):
            op.assembled_block(e, slot)
