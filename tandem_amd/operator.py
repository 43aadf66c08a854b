"""Host-side mirror of tandem's operator interface over the C ABI.

`DGOperator` mirrors tandem's AbstractDGOperator / DGOperator<LocalOperator>
(src/form/AbstractDGOperator.h:18-50, src/form/DGOperator.h:26-400): block_size(),
num_local_elements(), number_of_local_dofs(), apply(x, y), rhs(...).  `Elasticity` and `Poisson`
mirror the local operators' constructors (app/localoperator/Elasticity.h:42-44, Poisson.h), with
the coefficient functionals replaced by arrays evaluated at the operator's volume points.

Device vectors are torch tensors on the GPU (torch is used for device memory and streams only; all
arithmetic runs in the library's HIP kernels).
"""
import ctypes as C
from math import comb

import numpy as np

from . import _lib
from ._lib import check, lib, ptr

BC_NONE, BC_NATURAL, BC_FAULT, BC_DIRICHLET = 0, 1, 3, 5
ELASTICITY, POISSON = 0, 1


class Mesh:
    """Simplex mesh input (vertices, elements, boundary facets with BCs)."""

    def __init__(self, vertices, elements, boundary_facets=None, boundary_bc=None):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.dim = self.vertices.shape[1]
        self.elements = np.sort(np.ascontiguousarray(elements, dtype=np.int64), axis=1)
        if boundary_facets is None:
            boundary_facets = np.zeros((0, self.dim), dtype=np.int64)
            boundary_bc = np.zeros(0, dtype=np.int32)
        self.boundary_facets = np.ascontiguousarray(np.sort(np.asarray(boundary_facets, dtype=np.int64), axis=1))
        self.boundary_bc = np.ascontiguousarray(boundary_bc, dtype=np.int32)

    @property
    def num_elements(self):
        return self.elements.shape[0]

    @staticmethod
    def box(dim, n, scheme=0, bc=BC_DIRICHLET):
        """[0,1]^dim with n cells per direction (GenMesh tessellation, or Kuhn cubes for scheme 1)."""
        n = np.ascontiguousarray(np.broadcast_to(np.asarray(n, dtype=np.int64), (dim,)))
        bcs = np.ascontiguousarray(np.broadcast_to(np.asarray(bc, dtype=np.int32), (2 * dim,)))
        sizes = np.zeros(3, dtype=np.int64)
        check(lib().tdg_box_mesh(dim, n, scheme, bcs, sizes, None, None, None, None))
        nv, ne, nb = (int(s) for s in sizes)
        v = np.zeros((nv, dim))
        e = np.zeros((ne, dim + 1), dtype=np.int64)
        bf = np.zeros((nb, dim), dtype=np.int64)
        bb = np.zeros(nb, dtype=np.int32)
        check(lib().tdg_box_mesh(dim, n, scheme, bcs, sizes, ptr(v), ptr(e), ptr(bf), ptr(bb)))
        return Mesh(v, e, bf, bb)

    @staticmethod
    def union(meshes):
        """Disjoint union of meshes (vertex ids offset, elements kept in order): one operator over it
        holds every mesh; returns the mesh and the element offsets (len(meshes) + 1)."""
        assert len({m.dim for m in meshes}) == 1
        voff = np.cumsum([0] + [m.vertices.shape[0] for m in meshes])
        eoff = np.cumsum([0] + [m.num_elements for m in meshes])
        v = np.concatenate([m.vertices for m in meshes])
        e = np.concatenate([m.elements + voff[i] for i, m in enumerate(meshes)])
        bf = np.concatenate([m.boundary_facets + voff[i] for i, m in enumerate(meshes)])
        bb = np.concatenate([m.boundary_bc for m in meshes])
        return Mesh(v, e, bf, bb), eoff

    def affine_nodes(self, degree):
        """Warp-and-blend geometry nodes of the affine map, (elements, nbf, dim)."""
        ref = reference_nodes(self.dim, degree)
        v = self.vertices[self.elements]  # (E, D+1, D)
        lam0 = 1.0 - ref.sum(axis=1)
        return lam0[None, :, None] * v[:, None, 0, :] + np.einsum("pd,edi->epi", ref, v[:, 1:, :])


def reference_nodes(dim, degree):
    out = np.zeros(comb(degree + dim, dim) * dim)
    check(lib().tdg_reference_nodes(dim, degree, out))
    return out.reshape(-1, dim)


def interpolation_matrix(dim, to_degree, from_degree):
    """Nodal interpolation matrix (n_to x n_from) from degree from_degree to to_degree."""
    nt, nf = comb(to_degree + dim, dim), comb(from_degree + dim, dim)
    out = np.zeros(nt * nf)
    check(lib().tdg_interpolation_matrix(dim, to_degree, from_degree, out))
    return out.reshape(nt, nf)


def interpolate(I, x, num_quantities, transpose=False, y=None):
    """Per element and quantity y = I x (or I^T x) on the GPU; I is a (n_to, n_from) float64 GPU
    tensor and x holds elements x quantities x (n_from or n_to) values."""
    torch = _torch()
    nt, nf = I.shape
    n_in, n_out = (nt, nf) if transpose else (nf, nt)
    ip = _check(I, "interpolate: I")
    if x.numel() % (n_in * num_quantities) != 0:
        raise ValueError("interpolate: x holds no whole number of element blocks")
    xp = _check(x, "interpolate: x", like=I)
    E = x.numel() // (n_in * num_quantities)
    y = torch.empty(E * num_quantities * n_out, dtype=x.dtype, device=x.device) if y is None else y
    yp = _check(y, "interpolate: y", numel=E * num_quantities * n_out, like=I)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    check(lib().tdg_interpolate(ip, nt, nf, num_quantities, E, 1 if transpose else 0, xp, yp, C.c_void_p(stream)))
    return y


def _torch():
    import torch
    return torch


_DTYPES = {"f64": ("float64",), "f32": ("float32",), "any": ("float64", "float32")}


def _check(t, what, dtype="f64", numel=None, min_numel=None, cols=None, like=None):
    """Validate a device tensor before its pointer goes to the library: dtype, contiguity, size
    (exact numel, at least min_numel, or a (ncols, >= width) block given as cols=(ncols, width)),
    GPU residency and the device of `like`.  Raises TypeError / ValueError; returns the pointer."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s: a torch tensor is required" % what)
    if str(t.dtype).replace("torch.", "") not in _DTYPES[dtype]:
        raise TypeError("%s: dtype %s, expected %s" % (what, t.dtype, " or ".join(_DTYPES[dtype])))
    if like is not None and t.dtype != like.dtype:
        raise TypeError("%s: dtype %s differs from %s" % (what, t.dtype, like.dtype))
    if not t.is_contiguous():
        raise ValueError("%s: a contiguous tensor is required" % what)
    if numel is not None and t.numel() != numel:
        raise ValueError("%s: %d entries, expected %d" % (what, t.numel(), numel))
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError("%s: %d entries, expected at least %d" % (what, t.numel(), min_numel))
    if cols is not None:
        nc, width = cols
        if t.dim() != 2 or (nc is not None and t.shape[0] != nc) or t.shape[1] < width:
            raise ValueError("%s: shape %s, expected (%s, >= %d)" % (what, tuple(t.shape), nc if nc else "ncols",
                                                                      width))
    if not t.is_cuda:
        raise ValueError("%s: a GPU tensor is required" % what)
    if like is not None and t.device != like.device:
        raise ValueError("%s: on %s, expected %s" % (what, t.device, like.device))
    return C.c_void_p(t.data_ptr())


# kernel variants, indexed by TDG_VARIANT_*
VARIANTS = ["auto", "points", "mfma-xcd", "mfma-element", "assembled"]
# launch shapes of the two-kernel apply per precision (TDG_SPLIT_SHAPES) and workgroup orders
# (TDG_SPLIT_ORDERS: linear, per XCD, per XCD persistent)
SPLIT_SHAPES = 4
SPLIT_ORDERS = 3
# preconditioners, indexed by TDG_PRECOND_*: block Jacobi with FP64 blocks for FP64 solves, or with
# the FP32-stored blocks for every solve
_PRECONDS = {None: 0, "block-jacobi": 1, "block-jacobi-f32": 2}


class DGOperator:
    """Matrix-free SIPG operator y = A x on one GPU."""

    def __init__(self, mesh, kind, degree, geometry_nodes=None):
        self.mesh = mesh
        self.kind = kind
        self.degree = degree
        nodes = None
        if geometry_nodes is not None:
            nodes = np.ascontiguousarray(geometry_nodes, dtype=np.float64)
            assert nodes.size == mesh.num_elements * comb(degree + mesh.dim, mesh.dim) * mesh.dim
        desc = _lib.MeshDesc(mesh.dim, degree, mesh.vertices.shape[0], mesh.vertices.ctypes.data,
                             mesh.num_elements, mesh.elements.ctypes.data, mesh.boundary_bc.shape[0],
                             mesh.boundary_facets.ctypes.data if mesh.boundary_bc.shape[0] else None,
                             mesh.boundary_bc.ctypes.data if mesh.boundary_bc.shape[0] else None,
                             nodes.ctypes.data if nodes is not None else None)
        h = C.c_void_p()
        check(lib().tdg_create(C.byref(desc), kind, C.byref(h)))
        self._h = h
        s = np.zeros(7, dtype=np.int64)
        check(lib().tdg_sizes(h, s))
        self.bs, self.E, self.n, self.F, self.Nq, self.nq, self.dim = (int(x) for x in s)
        self.Q = self.dim if kind == ELASTICITY else 1
        self.Nbf = self.bs // self.Q
        self._material = False

    # --- AbstractDGOperator interface
    def block_size(self):
        return self.bs

    def num_local_elements(self):
        return self.E

    def number_of_local_dofs(self):
        return self.n

    # --- coefficient plumbing
    def volume_points(self):
        out = np.zeros(self.E * self.Nq * self.dim)
        check(lib().tdg_volume_points(self._h, out))
        return out.reshape(self.E, self.Nq, self.dim)

    def facet_points(self):
        out = np.zeros(self.F * self.nq * self.dim)
        check(lib().tdg_facet_points(self._h, out))
        return out.reshape(self.F, self.nq, self.dim)

    def facets(self):
        out = np.zeros(self.F * 5, dtype=np.int64)
        check(lib().tdg_facets(self._h, out))
        return out.reshape(self.F, 5)

    def set_material(self, c0_Q, c1_Q=None):
        c0 = np.ascontiguousarray(c0_Q, dtype=np.float64).reshape(self.E * self.Nq)
        c1 = None if c1_Q is None else np.ascontiguousarray(c1_Q, dtype=np.float64).reshape(self.E * self.Nq)
        check(lib().tdg_set_material(self._h, ptr(c0), ptr(c1)))
        self._material = True

    def set_variant(self, variant):
        """"auto" (two-kernel matrix-core apply where compiled: facet kernel + element kernel),
        "points" (thread-per-point kernel), "mfma-element" (element-centric matrix-core kernel) or
        "mfma-xcd" (the same, workgroups renumbered so each XCD takes one element range) or
        "assembled" (the assembled block rows, assembled on first use)."""
        check(lib().tdg_set_variant(self._h, VARIANTS.index(variant)))

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream (or None for the default stream).  The library's calls are
        asynchronous on that stream; make it torch's current stream as well (torch.cuda.set_stream
        or a `with torch.cuda.stream(...)` block), so that the tensors passed in are produced, and
        their memory recycled by torch's allocator, in the same stream order."""
        check(lib().tdg_set_stream(self._h, None if stream is None else C.c_void_p(stream.cuda_stream)))

    # --- apply
    def apply(self, x, y):
        """y = A x on GPU tensors (float64 or float32, contiguous, number_of_local_dofs() long)."""
        torch = _torch()
        xp = _check(x, "apply: x", "any", min_numel=self.n)
        yp = _check(y, "apply: y", "any", min_numel=self.n, like=x)
        f = lib().tdg_mult if x.dtype == torch.float64 else lib().tdg_mult_f32
        check(f(self._h, xp, yp))
        return y

    def apply_batched(self, X, Y):
        """Y[:, c] = A X[:, c]; X, Y of shape (ncols, ld) row-major (each column contiguous)."""
        torch = _torch()
        xp = _check(X, "apply_batched: X", "any", cols=(None, self.n))
        yp = _check(Y, "apply_batched: Y", "any", cols=(X.shape[0], self.n), like=X)
        if Y.shape != X.shape:
            raise ValueError("apply_batched: X and Y shapes differ")
        f = lib().tdg_mult_batched if X.dtype == torch.float64 else lib().tdg_mult_batched_f32
        check(f(self._h, xp, yp, X.shape[0], X.shape[1]))
        return Y

    def upload(self):
        """Make the operator device-resident now (otherwise the first apply does it)."""
        check(lib().tdg_upload(self._h))

    def solve(self, b, x, rtol=1e-10, mode="fp64", inner_rtol=1e-4, max_iter=200000, max_outer=60,
              check_every=10, reliable_delta=0.1, precond=None, graphs=True):
        """Solve A x = b (float64 GPU tensors).  mode "fp64": FP64 CG; "fp32-ir": FP32 CG solves
        restarted inside FP64 iterative refinement; "fp32-ru": FP32 CG with FP64 residual
        replacement (refinement that keeps the Krylov direction).  precond: None or
        "block-jacobi".  graphs: replay the iterations between residual checks as a captured HIP
        graph when a stream is set.  Returns the statistics."""
        bp = _check(b, "solve: b", numel=self.n)
        xp = _check(x, "solve: x", numel=self.n, like=b)
        opts = _lib.SolveOpts(rtol, max_iter, {"fp64": 0, "fp32-ir": 1, "fp32-ru": 2}[mode], inner_rtol, max_outer,
                              check_every, reliable_delta, _PRECONDS[precond], int(graphs))
        res = _lib.SolveResult()
        check(lib().tdg_solve(self._h, bp, xp, C.byref(opts), C.byref(res)))
        return {k: getattr(res, k) for k, _ in _lib.SolveResult._fields_}

    def solve_multi(self, B, X, rtol=1e-10, mode="fp64", max_iter=200000, check_every=10, reliable_delta=0.1,
                    precond=None, graphs=True):
        """Solve A X[c] = B[c] for the rows c of (ncols, ld) float64 GPU tensors together (one batched
        apply per iteration).  mode "fp64" or "fp32-ru".  Returns the statistics with per-column
        iterations and relative residuals."""
        bp = _check(B, "solve_multi: B", cols=(None, self.n))
        xp = _check(X, "solve_multi: X", cols=(B.shape[0], self.n), like=B)
        if X.shape != B.shape:
            raise ValueError("solve_multi: B and X shapes differ")
        nc, ld = B.shape
        opts = _lib.SolveOpts(rtol, max_iter, {"fp64": 0, "fp32-ru": 2}[mode], 1e-4, 1, check_every, reliable_delta,
                              _PRECONDS[precond], int(graphs))
        res = _lib.SolveResult()
        its = np.zeros(nc, dtype=np.int32)
        rel = np.zeros(nc)
        check(lib().tdg_solve_multi(self._h, bp, xp, nc, ld, C.byref(opts), C.byref(res), its, rel))
        out = {k: getattr(res, k) for k, _ in _lib.SolveResult._fields_}
        out["column_iterations"] = its
        out["column_rel_residual"] = rel
        return out

    def autotune(self, ncols=1, precision="fp64"):
        """Time the applicable kernel variants on ncols columns, keep the fastest; returns the chosen
        variant name and the milliseconds per apply of each."""
        chosen = C.c_int()
        ms = np.zeros(len(VARIANTS))
        check(lib().tdg_autotune(self._h, ncols, 8 if precision == "fp64" else 4, C.byref(chosen), ms))
        return VARIANTS[chosen.value], {VARIANTS[v]: float(ms[v]) for v in range(len(VARIANTS)) if ms[v] >= 0}

    def autotune_shapes(self, ncols=1, precision="fp64"):
        """Time the launch shapes of the two-kernel apply (shape s, workgroup order x: 0 linear, 1 per
        XCD, 2 per XCD with a persistent launch) on ncols columns, keep the fastest for that
        precision; returns ((s, x), {(s, x): ms})."""
        shape, xcd = C.c_int(), C.c_int()
        n = SPLIT_ORDERS * SPLIT_SHAPES
        ms = np.zeros(n)
        check(lib().tdg_autotune_shapes(self._h, ncols, 8 if precision == "fp64" else 4, C.byref(shape),
                                        C.byref(xcd), ms))
        return (shape.value, xcd.value), {(c // SPLIT_ORDERS, c % SPLIT_ORDERS): float(ms[c]) for c in range(n)
                                          if ms[c] >= 0}

    def set_shape(self, precision, shape, xcd=0):
        """Launch shape of the two-kernel apply for "fp64" or "fp32"."""
        check(lib().tdg_set_shape(self._h, 8 if precision == "fp64" else 4, shape, xcd))

    def solve_segments(self, b, x, offsets, rtol=1e-10, mode="fp64", max_iter=200000, check_every=10,
                       reliable_delta=0.1, precond=None, graphs=True):
        """Independent meshes in one operator: solve each segment [offsets[c], offsets[c+1]) of the
        float64 GPU vectors b, x as its own system, all in lockstep (one apply per iteration)."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        if off.ndim != 1 or len(off) < 2 or off[0] < 0 or off[-1] > self.n or np.any(np.diff(off) <= 0):
            raise ValueError("solve_segments: offsets must increase strictly within [0, n]")
        bp = _check(b, "solve_segments: b", numel=self.n)
        xp = _check(x, "solve_segments: x", numel=self.n, like=b)
        ns = len(off) - 1
        opts = _lib.SolveOpts(rtol, max_iter, {"fp64": 0, "fp32-ru": 2}[mode], 1e-4, 1, check_every, reliable_delta,
                              _PRECONDS[precond], int(graphs))
        res = _lib.SolveResult()
        its = np.zeros(ns, dtype=np.int32)
        rel = np.zeros(ns)
        check(lib().tdg_solve_segments(self._h, bp, xp, ns, off, C.byref(opts), C.byref(res), its, rel))
        out = {k: getattr(res, k) for k, _ in _lib.SolveResult._fields_}
        out["segment_iterations"] = its
        out["segment_rel_residual"] = rel
        return out

    def rhs(self, force_Q=None, facet_q=None):
        """FP64 right-hand side on the host.  force: (E, Q, Nq); facet data (F, Q, nq): Dirichlet
        values on Dirichlet facets, slip on Fault facets."""
        b = np.zeros(self.n)
        f = None if force_Q is None else np.ascontiguousarray(force_Q, dtype=np.float64)
        d = None if facet_q is None else np.ascontiguousarray(facet_q, dtype=np.float64)
        check(lib().tdg_rhs(self._h, ptr(f), ptr(d), b))
        return b

    def rhs_device(self, b, force=None, facet=None):
        """FP64 right-hand side on the GPU into b (elasticity).  Tensors on the GPU: force (E*Q*Nq)
        or (ncols, >= E*Q*Nq), facet data (F*Q*nq) or (ncols, ...), b (n) or (ncols, >= n)."""
        torch = _torch()
        ncols = 1 if b.dim() == 1 else b.shape[0]

        def view(t, width, what):
            if t is None:
                return None, 0
            if ncols == 1:
                return _check(t, "rhs_device: " + what, min_numel=width, like=b), width
            return _check(t, "rhs_device: " + what, cols=(ncols, width), like=b), t.shape[1]

        _check(b, "rhs_device: b")
        fp, lf = view(force, self.E * self.Q * self.Nq, "force")
        dp, ld = view(facet, self.F * self.Q * self.nq, "facet")
        bp, lb = view(b, self.n, "b")
        check(lib().tdg_rhs_device(self._h, fp, dp, bp, ncols, lf, ld, lb))
        return b

    def traction(self, x, facet=None, t=None):
        """Facet traction (F, Q, nq) of the GPU vector x (float64) with facet data (may be None)."""
        torch = _torch()
        width = self.F * self.Q * self.nq
        xp = _check(x, "traction: x", numel=self.n)
        if t is None:
            t = torch.empty(width, dtype=torch.float64, device=x.device)
        tp = _check(t, "traction: t", numel=width, like=x)
        fp = None if facet is None else _check(facet, "traction: facet", numel=width, like=x)
        check(lib().tdg_traction(self._h, xp, fp, tp))
        return t.view(self.F, self.Q, self.nq)

    def block_jacobi_build(self):
        """Build the element-block Jacobi preconditioner; returns the number of element colours."""
        c = C.c_int()
        check(lib().tdg_block_jacobi_build(self._h, C.byref(c)))
        return c.value

    def block_jacobi_apply(self, r, z):
        """z = D^-1 r (float64 GPU tensors)."""
        rp = _check(r, "block_jacobi_apply: r", numel=self.n)
        zp = _check(z, "block_jacobi_apply: z", numel=self.n, like=r)
        check(lib().tdg_block_jacobi_apply(self._h, rp, zp))
        return z

    def block_jacobi_apply_batched(self, R, Z):
        """Z[c] = D^-1 R[c] for the rows c of (ncols, ld) float64 GPU tensors."""
        rp = _check(R, "block_jacobi_apply_batched: R", cols=(None, self.n))
        zp = _check(Z, "block_jacobi_apply_batched: Z", cols=(R.shape[0], self.n), like=R)
        if Z.shape != R.shape:
            raise ValueError("block_jacobi_apply_batched: R and Z shapes differ")
        check(lib().tdg_block_jacobi_apply_batched(self._h, rp, zp, R.shape[0], R.shape[1]))
        return Z

    def block_jacobi_inverse(self, element):
        out = np.zeros(self.bs * self.bs)
        check(lib().tdg_block_jacobi_inverse(self._h, element, out))
        return out.reshape(self.bs, self.bs, order="F")

    def assemble(self):
        """Assembled block rows on the device (tandem's PetscDGMatrix, DGOperator::assemble), probed
        from the FP64 matrix-free apply; returns the number of probe colours.  The "assembled"
        variant multiplies by them."""
        c = C.c_int()
        check(lib().tdg_assemble(self._h, C.byref(c)))
        return c.value

    def assembled_block(self, element, slot):
        """Block `slot` (0: own, 1 + f: face-f neighbour) of an element's assembled block row."""
        out = np.zeros(self.bs * self.bs)
        check(lib().tdg_assembled_block(self._h, element, slot, out))
        return out.reshape(self.bs, self.bs)

    def material_gradient(self, u, psi, g_lam=None, g_mu=None):
        """d/d(lam_Q, mu_Q) of psi^T A u for GPU vectors u, psi (float64 or float32); returns two
        (E, Nq) tensors."""
        torch = _torch()
        up = _check(u, "material_gradient: u", "any", numel=self.n)
        pp = _check(psi, "material_gradient: psi", "any", numel=self.n, like=u)
        if g_lam is None:
            g_lam = torch.empty(self.E * self.Nq, dtype=u.dtype, device=u.device)
        if g_mu is None:
            g_mu = torch.empty_like(g_lam)
        gl = _check(g_lam, "material_gradient: g_lam", "any", numel=self.E * self.Nq, like=u)
        gm = _check(g_mu, "material_gradient: g_mu", "any", numel=self.E * self.Nq, like=u)
        f = lib().tdg_material_gradient if u.dtype == torch.float64 else lib().tdg_material_gradient_f32
        check(f(self._h, up, pp, gl, gm))
        return g_lam.view(self.E, self.Nq), g_mu.view(self.E, self.Nq)

    def fault_setup(self, up, ref_normal):
        """Fault adapter on the Fault facets; returns (number of fault facets, nbf of the fault space)."""
        sizes = np.zeros(2, dtype=np.int64)
        check(lib().tdg_fault_setup(self._h, np.ascontiguousarray(up, dtype=np.float64),
                                    np.ascontiguousarray(ref_normal, dtype=np.float64), sizes))
        self.nfault, self.fault_nbf = int(sizes[0]), int(sizes[1])
        return self.nfault, self.fault_nbf

    def fault_slip(self, S, F=None):
        """Facet data (F*D*nq per column) of fault-basis slips S: (nfault*nbf*(D-1)) or (ncols, ...)
        float64 GPU tensors."""
        torch = _torch()
        ncols = 1 if S.dim() == 1 else S.shape[0]
        width = self.F * self.dim * self.nq
        ws = self.nfault * self.fault_nbf * (self.dim - 1)
        if F is None:
            F = torch.empty((ncols, width) if S.dim() == 2 else (width,), dtype=torch.float64, device=S.device)
        if S.dim() == 2:
            sp = _check(S, "fault_slip: S", cols=(None, ws))
            fp = _check(F, "fault_slip: F", cols=(ncols, width), like=S)
        else:
            sp = _check(S, "fault_slip: S", numel=ws)
            fp = _check(F, "fault_slip: F", numel=width, like=S)
        ld_s = S.shape[-1] if S.dim() == 2 else 0
        ld_f = F.shape[-1] if F.dim() == 2 else 0
        check(lib().tdg_fault_slip(self._h, sp, fp, ncols, ld_s, ld_f))
        return F

    def fault_traction(self, X, Fd=None, T=None):
        """Fault-basis traction (nfault*nbf*D per column) of solutions X (n) or (ncols, >= n) with
        facet data Fd (same column layout, may be None)."""
        torch = _torch()
        ncols = 1 if X.dim() == 1 else X.shape[0]
        width = self.nfault * self.fault_nbf * self.dim
        wf = self.F * self.dim * self.nq
        if T is None:
            T = torch.empty((ncols, width) if X.dim() == 2 else (width,), dtype=torch.float64, device=X.device)
        if X.dim() == 2:
            xp = _check(X, "fault_traction: X", cols=(None, self.n))
            tp = _check(T, "fault_traction: T", cols=(ncols, width), like=X)
            fp = None if Fd is None else _check(Fd, "fault_traction: Fd", cols=(ncols, wf), like=X)
        else:
            xp = _check(X, "fault_traction: X", numel=self.n)
            tp = _check(T, "fault_traction: T", numel=width, like=X)
            fp = None if Fd is None else _check(Fd, "fault_traction: Fd", numel=wf, like=X)
        ld_x = X.shape[-1] if X.dim() == 2 else 0
        ld_f = Fd.shape[-1] if (Fd is not None and Fd.dim() == 2) else 0
        ld_t = T.shape[-1] if T.dim() == 2 else 0
        check(lib().tdg_fault_traction(self._h, xp, fp, tp, ncols, ld_x, ld_f, ld_t))
        return T

    def inverse_mass(self, x, y=None):
        """y = M^-1 x (weight-adjusted inverse mass with the reference's negative sign) for a float64
        GPU vector x."""
        torch = _torch()
        xp = _check(x, "inverse_mass: x", numel=self.n)
        y = torch.empty_like(x) if y is None else y
        yp = _check(y, "inverse_mass: y", numel=self.n, like=x)
        check(lib().tdg_inverse_mass(self._h, xp, yp))
        return y

    def wave_rhs(self, x, facet=None, y=None):
        """y = M^-1 (A x - b_facet(facet)) for a float64 GPU vector x and optional GPU facet data
        (facets x quantities x facet points)."""
        torch = _torch()
        xp = _check(x, "wave_rhs: x", numel=self.n)
        y = torch.empty_like(x) if y is None else y
        yp = _check(y, "wave_rhs: y", numel=self.n, like=x)
        fp = None if facet is None else _check(facet, "wave_rhs: facet", numel=self.F * self.Q * self.nq, like=x)
        check(lib().tdg_wave_rhs(self._h, xp, fp, yp))
        return y

    def apply_counts(self, precision_bytes=8):
        b, f = C.c_double(), C.c_double()
        check(lib().tdg_apply_counts(self._h, precision_bytes, C.byref(b), C.byref(f)))
        return b.value, f.value

    def host_array(self, name):
        n = C.c_int64()
        check(lib().tdg_host_array(self._h, name.encode(), None, C.byref(n)))
        out = np.zeros(n.value)
        check(lib().tdg_host_array(self._h, name.encode(), C.c_void_p(out.ctypes.data), C.byref(n)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tdg_destroy(h)
            self._h = None


class PMultigrid:
    """p-multigrid hierarchy over DG operators of decreasing degree on one mesh (levels[0] the fine
    one; the coarse operators are built by the caller with the same material).  Mirrors the PCMG
    that PetscLinearSolver::setup_mg (app/common/PetscLinearSolver.cpp:53-114) builds: nodal
    interpolation between levels, Chebyshev / block-Jacobi smoothing, a block-Jacobi CG coarse
    solve to coarse_rtol.  The operators must outlive the hierarchy (it keeps references)."""

    def __init__(self, levels, smooth=None, eig_lo=None, eig_hi=None, eig_iters=None, coarse_rtol=None,
                 coarse_max_iter=None, galerkin_penalty=None, coarse_fp32=None, vcycle_fp32=None,
                 coarse_assembled=None, coarse_amg=None):
        o = _lib.PMGOpts()
        lib().tdg_pmg_default_opts(C.byref(o))
        for k, v in (("smooth", smooth), ("eig_lo", eig_lo), ("eig_hi", eig_hi), ("eig_iters", eig_iters),
                     ("coarse_rtol", coarse_rtol), ("coarse_max_iter", coarse_max_iter),
                     ("galerkin_penalty", galerkin_penalty), ("coarse_fp32", coarse_fp32),
                     ("vcycle_fp32", vcycle_fp32), ("coarse_assembled", coarse_assembled),
                     ("coarse_amg", coarse_amg)):
            if v is not None:
                setattr(o, k, type(getattr(o, k))(v))
        self.options = {k: getattr(o, k) for k, _ in _lib.PMGOpts._fields_}
        self.levels = list(levels)
        arr = (C.c_void_p * len(self.levels))(*[op._h for op in self.levels])
        h = C.c_void_p()
        check(lib().tdg_pmg_create(arr, len(self.levels), C.byref(o), C.byref(h)))
        self._h = h
        self.n = self.levels[0].n

    def setup(self):
        """Block-Jacobi inverses and eigenvalue estimates; returns lambda_max per smoothed level."""
        lm = np.zeros(max(len(self.levels) - 1, 1))
        check(lib().tdg_pmg_setup(self._h, lm))
        return lm[: len(self.levels) - 1]

    def vcycle(self, R, Z):
        """Z = M^-1 R for (ncols, ld) float64 GPU tensors (one V-cycle)."""
        rp = _check(R, "vcycle: R", cols=(None, self.n))
        zp = _check(Z, "vcycle: Z", cols=(R.shape[0], self.n), like=R)
        if R.shape != Z.shape:
            raise ValueError("vcycle: R and Z shapes differ")
        check(lib().tdg_pmg_vcycle(self._h, rp, zp, R.shape[0], R.shape[1]))

    def solve_multi(self, B, X, rtol=1e-10, max_iter=1000):
        """Solve A X[c] = B[c] (rows of (ncols, ld) float64 GPU tensors) by flexible CG with one
        V-cycle per iteration.  Returns the statistics with per-column iterations and relative
        residuals; outer_iterations = coarse CG iterations over all V-cycles."""
        bp = _check(B, "solve_multi: B", cols=(None, self.n))
        xp = _check(X, "solve_multi: X", cols=(B.shape[0], self.n), like=B)
        if X.shape != B.shape:
            raise ValueError("solve_multi: B and X shapes differ")
        nc, ld = B.shape
        opts = _lib.SolveOpts(rtol, max_iter, 0, 1e-4, 1, 1, 0.1, 0, 0)
        res = _lib.SolveResult()
        its = np.zeros(nc, dtype=np.int32)
        rel = np.zeros(nc)
        check(lib().tdg_solve_multi_pmg(self._h, bp, xp, nc, ld, C.byref(opts), C.byref(res), its, rel))
        out = {k: getattr(res, k) for k, _ in _lib.SolveResult._fields_}
        out["column_iterations"] = its
        out["column_rel_residual"] = rel
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tdg_pmg_destroy(h)
            self._h = None


class Elasticity(DGOperator):
    """Linear elasticity (tandem Elasticity): lam_Q, mu_Q at the volume points, or callables f(x)."""

    def __init__(self, mesh, degree, lam=None, mu=None, geometry_nodes=None):
        super().__init__(mesh, ELASTICITY, degree, geometry_nodes)
        if lam is not None:
            X = self.volume_points()
            lq = lam(X) if callable(lam) else lam
            mq = mu(X) if callable(mu) else mu
            self.set_material(np.broadcast_to(lq, X.shape[:2]), np.broadcast_to(mq, X.shape[:2]))


class Poisson(DGOperator):
    """Scalar diffusion (tandem Poisson, modal Dubiner basis): K_Q at the volume points or f(x)."""

    def __init__(self, mesh, degree, K=None, geometry_nodes=None):
        super().__init__(mesh, POISSON, degree, geometry_nodes)
        if K is not None:
            X = self.volume_points()
            kq = K(X) if callable(K) else K
            self.set_material(np.broadcast_to(kq, X.shape[:2]))
