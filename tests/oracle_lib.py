"""ctypes view of the CPU FP64 oracle (oracle/_build/liboracle.so) for the test suite.

The oracle is the checker: tests compare the HIP path against it; nothing in tandem_amd uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
RULES = os.path.join(ROOT, "tandem_amd", "data", "simplex_rules.txt")

BC_NONE, BC_NATURAL, BC_FAULT, BC_DIRICHLET = 0, 1, 3, 5

_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
up32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
up64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    L = C.CDLL(LIB_PATH)
    L.oracle_last_error.restype = C.c_char_p
    L.oracle_load_rules.argtypes = [C.c_char_p]
    L.oracle_jacobi_p.restype = C.c_double
    L.oracle_jacobi_p.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_double]
    L.oracle_sf_jacobi_p.restype = C.c_double
    L.oracle_sf_jacobi_p.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_double, C.c_double]
    L.oracle_dubiner.restype = C.c_double
    L.oracle_dubiner.argtypes = [C.c_uint, up32, dp]
    L.oracle_grad_dubiner.argtypes = [C.c_uint, up32, dp, dp]
    L.oracle_all_integer_sums.argtypes = [C.c_uint, C.c_uint, C.c_void_p]
    L.oracle_gauss_jacobi.argtypes = [C.c_uint, C.c_uint, C.c_uint, dp, dp]
    L.oracle_lgl_points.argtypes = [C.c_uint, dp]
    L.oracle_rule.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    L.oracle_wab_nodes.argtypes = [C.c_uint, C.c_uint, C.c_void_p]
    L.oracle_ref_mass.argtypes = [C.c_uint, C.c_uint, C.c_int, C.c_int, dp]
    L.oracle_eval_basis.argtypes = [C.c_uint, C.c_uint, C.c_int, C.c_int, dp, C.c_void_p, C.c_void_p]
    L.oracle_box_mesh.restype = C.c_void_p
    L.oracle_box_mesh.argtypes = [C.c_uint, up64, C.c_int, ip32]
    L.oracle_mesh_new.restype = C.c_void_p
    L.oracle_mesh_new.argtypes = [C.c_int, C.c_int64, dp, C.c_int64, ip64, C.c_int64, ip64, ip32]
    L.oracle_mesh_sizes.argtypes = [C.c_void_p, ip64]
    L.oracle_mesh_arrays.argtypes = [C.c_void_p, dp, ip64, ip64, ip32]
    L.oracle_mesh_free.argtypes = [C.c_void_p]
    L.oracle_affine_nodes.argtypes = [C.c_void_p, C.c_uint, dp]
    L.oracle_geometry_coords.argtypes = [C.c_void_p, C.c_uint, C.c_void_p, ip64, C.c_void_p, C.c_void_p]
    L.oracle_op_new.restype = C.c_void_p
    L.oracle_op_new.argtypes = [C.c_int, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    L.oracle_op_free.argtypes = [C.c_void_p]
    L.oracle_op_info.argtypes = [C.c_void_p, ip64]
    L.oracle_op_apply.argtypes = [C.c_void_p, dp, dp, C.c_int]
    L.oracle_op_assemble.argtypes = [C.c_void_p, dp]
    L.oracle_op_rhs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, dp]
    L.oracle_op_traction.argtypes = [C.c_void_p, C.c_int64, dp, C.c_void_p, dp]
    L.oracle_op_fault_setup.argtypes = [C.c_void_p, dp, dp, ip64]
    L.oracle_op_fault_traction.argtypes = [C.c_void_p, C.c_int64, dp, dp]
    L.oracle_op_fault_slip.argtypes = [C.c_void_p, C.c_int64, dp, dp]
    L.oracle_op_fault_arrays.argtypes = [C.c_void_p, ip64, dp, dp, dp]
    L.oracle_op_material_gradient.argtypes = [C.c_void_p, dp, dp, dp, dp]
    L.oracle_op_wave_rhs.argtypes = [C.c_void_p, dp, C.c_void_p, dp]
    L.oracle_op_inverse_mass.argtypes = [C.c_void_p, dp, dp]
    L.oracle_nodal_interpolation.argtypes = [C.c_int, C.c_int, C.c_int, dp]
    L.oracle_op_facets.argtypes = [C.c_void_p, ip64]
    L.oracle_op_eval_volume.argtypes = [C.c_void_p, dp, dp, C.c_void_p]
    L.oracle_op_array.argtypes = [C.c_void_p, C.c_char_p, dp]
    L.oracle_op_array_size.restype = C.c_int64
    L.oracle_op_array_size.argtypes = [C.c_void_p, C.c_char_p]
    if L.oracle_load_rules(RULES.encode()) != 0:
        raise RuntimeError(L.oracle_last_error().decode())
    _lib = L
    return L


def _ptr(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)


def _err():
    raise RuntimeError("oracle: " + lib().oracle_last_error().decode())


def all_integer_sums(D, N):
    L = lib()
    n = L.oracle_all_integer_sums(D, N, None)
    out = np.zeros(n * D, dtype=np.uint32)
    L.oracle_all_integer_sums(D, N, out.ctypes.data_as(C.c_void_p))
    return out.reshape(n, D)


def dubiner(D, idx, xi):
    return lib().oracle_dubiner(D, np.asarray(idx, dtype=np.uint32), np.asarray(xi, dtype=np.float64))


def grad_dubiner(D, idx, xi):
    g = np.zeros(3)
    lib().oracle_grad_dubiner(D, np.asarray(idx, dtype=np.uint32), np.asarray(xi, dtype=np.float64), g)
    return g[:D]


def gauss_jacobi(n, a, b):
    p, w = np.zeros(n), np.zeros(n)
    lib().oracle_gauss_jacobi(n, a, b, p, w)
    return p, w


def lgl_points(n):
    p = np.zeros(n)
    lib().oracle_lgl_points(n, p)
    return p


def rule(D, order, family=0):
    """family: 0 auto, 1 tensor product, 2 Witherden-Vincent, 3 Jaskowiec-Sukumar."""
    L = lib()
    o = C.c_int(0)
    n = L.oracle_rule(family, D, order, None, None, C.byref(o))
    if n < 0:
        return None
    p, w = np.zeros(n * D), np.zeros(n)
    L.oracle_rule(family, D, order, p.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p), C.byref(o))
    return p.reshape(n, D), w, o.value


def wab_nodes(D, N):
    L = lib()
    n = L.oracle_wab_nodes(D, N, None)
    out = np.zeros(n * D)
    L.oracle_wab_nodes(D, N, out.ctypes.data_as(C.c_void_p))
    return out.reshape(n, D)


def ref_mass(D, N, modal, inverse=False):
    from math import comb
    nbf = comb(N + D, D)
    out = np.zeros(nbf * nbf)
    if lib().oracle_ref_mass(D, N, int(modal), int(inverse), out) != 0:
        _err()
    return out.reshape(nbf, nbf, order="F")


def eval_basis(D, N, modal, pts):
    """Returns E (nbf, np) and G (nbf, D, np) as Fortran-ordered arrays."""
    from math import comb
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    npts = pts.shape[0]
    nbf = comb(N + D, D)
    E = np.zeros(nbf * npts)
    G = np.zeros(nbf * D * npts)
    if lib().oracle_eval_basis(D, N, int(modal), npts, pts.ravel(), E.ctypes.data_as(C.c_void_p),
                               G.ctypes.data_as(C.c_void_p)) != 0:
        _err()
    return E.reshape(nbf, npts, order="F"), G.reshape(nbf, D, npts, order="F")


class Mesh:
    def __init__(self, handle):
        self.h = handle
        s = np.zeros(4, dtype=np.int64)
        lib().oracle_mesh_sizes(handle, s)
        self.D, nv, ne, nb = (int(x) for x in s)
        self.verts = np.zeros(nv * self.D)
        self.elems = np.zeros(ne * (self.D + 1), dtype=np.int64)
        self.bfacets = np.zeros(nb * self.D, dtype=np.int64)
        self.bbc = np.zeros(nb, dtype=np.int32)
        lib().oracle_mesh_arrays(handle, self.verts, self.elems, self.bfacets, self.bbc)
        self.verts = self.verts.reshape(nv, self.D)
        self.elems = self.elems.reshape(ne, self.D + 1)
        self.bfacets = self.bfacets.reshape(nb, self.D)

    @staticmethod
    def box(D, N, tess, bcs):
        h = lib().oracle_box_mesh(D, np.asarray(N, dtype=np.uint64), tess, np.asarray(bcs, dtype=np.int32))
        if not h:
            _err()
        return Mesh(h)

    @staticmethod
    def from_arrays(verts, elems, bfacets, bbc):
        verts = np.ascontiguousarray(verts, dtype=np.float64)
        D = verts.shape[1]
        elems = np.ascontiguousarray(elems, dtype=np.int64)
        bfacets = np.ascontiguousarray(bfacets, dtype=np.int64).reshape(-1, D)
        bbc = np.ascontiguousarray(bbc, dtype=np.int32)
        h = lib().oracle_mesh_new(D, verts.shape[0], verts.ravel(), elems.shape[0], elems.ravel(),
                                  bbc.shape[0], bfacets.ravel(), bbc)
        return Mesh(h)

    @property
    def num_elements(self):
        return self.elems.shape[0]

    def affine_nodes(self, N):
        from math import comb
        out = np.zeros(self.num_elements * comb(N + self.D, self.D) * self.D)
        lib().oracle_affine_nodes(self.h, N, out)
        return out.reshape(self.num_elements, -1, self.D)

    def coords(self, N, nodes=None):
        """(vol coords (E, Nq, D), facet coords (F, nq, D))."""
        s = np.zeros(3, dtype=np.int64)
        nd = None if nodes is None else np.ascontiguousarray(nodes, dtype=np.float64)
        if lib().oracle_geometry_coords(self.h, N, _ptr(nd), s, None, None) != 0:
            _err()
        Nq, nq, F = (int(x) for x in s)
        v = np.zeros(self.num_elements * Nq * self.D)
        f = np.zeros(F * nq * self.D)
        lib().oracle_geometry_coords(self.h, N, _ptr(nd), s, v.ctypes.data_as(C.c_void_p),
                                     f.ctypes.data_as(C.c_void_p))
        return v.reshape(self.num_elements, Nq, self.D), f.reshape(F, nq, self.D)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_mesh_free(self.h)
            self.h = None


class Operator:
    """Elasticity (kind 0; coefficients lam_Q, mu_Q at volume points) or Poisson (kind 1; K_Q)."""

    def __init__(self, kind, mesh, N, c0, c1=None, nodes=None):
        self.mesh = mesh
        self.kind = kind
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        c1 = None if c1 is None else np.ascontiguousarray(c1, dtype=np.float64)
        nd = None if nodes is None else np.ascontiguousarray(nodes, dtype=np.float64)
        self._keep = (c0, c1, nd)
        self.h = lib().oracle_op_new(kind, mesh.h, N, _ptr(nd), _ptr(c0), _ptr(c1))
        if not self.h:
            _err()
        info = np.zeros(6, dtype=np.int64)
        lib().oracle_op_info(self.h, info)
        self.bs, self.E, self.F, self.Nq, self.nq, self.D = (int(x) for x in info)
        self.n = self.bs * self.E

    def apply(self, x, nthreads=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        if lib().oracle_op_apply(self.h, x, y, nthreads) != 0:
            _err()
        return y

    def assemble(self):
        A = np.zeros(self.n * self.n)
        if lib().oracle_op_assemble(self.h, A) != 0:
            _err()
        return A.reshape(self.n, self.n, order="F")

    def rhs(self, force_Q=None, dirichlet_q=None):
        b = np.zeros(self.n)
        fq = None if force_Q is None else np.ascontiguousarray(force_Q, dtype=np.float64)
        dq = None if dirichlet_q is None else np.ascontiguousarray(dirichlet_q, dtype=np.float64)
        if lib().oracle_op_rhs(self.h, _ptr(fq), _ptr(dq), b) != 0:
            _err()
        return b

    def traction(self, fct, x, f_q=None):
        t = np.zeros(self.D * self.nq)
        fq = None if f_q is None else np.ascontiguousarray(f_q, dtype=np.float64)
        if lib().oracle_op_traction(self.h, fct, np.ascontiguousarray(x, dtype=np.float64), _ptr(fq), t) != 0:
            _err()
        return t.reshape(self.D, self.nq)

    def fault_setup(self, up, ref_normal):
        """Fault adapter on the BC::Fault facets; returns (facets, basis (nf, nq, D, D) with
        basis[m, q, i, j] = component i of vector j, nl (nf, nq), e_q (nq, nbf))."""
        sz = np.zeros(3, dtype=np.int64)
        if lib().oracle_op_fault_setup(self.h, np.asarray(up, dtype=np.float64), np.asarray(ref_normal, dtype=np.float64),
                                       sz) != 0:
            _err()
        nf, nbf, nq = (int(v) for v in sz)
        self.fault_nbf = nbf
        fac = np.zeros(nf, dtype=np.int64)
        basis = np.zeros(nf * nq * self.D * self.D)
        nl = np.zeros(nf * nq)
        e_q = np.zeros(nbf * nq)
        if lib().oracle_op_fault_arrays(self.h, fac, basis, nl, e_q) != 0:
            _err()
        return fac, basis.reshape(nf, nq, self.D, self.D), nl.reshape(nf, nq), e_q.reshape(nq, nbf)

    def fault_traction(self, m, t_q):
        """Fault-basis traction (nbf, D) of facet-point traction t_q (D, nq) on fault facet m."""
        out = np.zeros(self.fault_nbf * self.D)
        if lib().oracle_op_fault_traction(self.h, m, np.ascontiguousarray(t_q, dtype=np.float64), out) != 0:
            _err()
        return out.reshape(self.D, self.fault_nbf).T

    def fault_slip(self, m, s):
        """Facet-point slip (D, nq) of the fault-basis slip s (nbf, D-1) on fault facet m."""
        out = np.zeros(self.D * self.nq)
        if lib().oracle_op_fault_slip(self.h, m, np.ascontiguousarray(np.asarray(s).T, dtype=np.float64).ravel(),
                                      out) != 0:
            _err()
        return out.reshape(self.D, self.nq)

    def material_gradient(self, u, psi):
        """d/d(lam_Q, mu_Q) of psi^T A u, each (E, Nq)."""
        gl = np.zeros(self.E * self.Nq)
        gm = np.zeros(self.E * self.Nq)
        if lib().oracle_op_material_gradient(self.h, np.ascontiguousarray(u, dtype=np.float64),
                                             np.ascontiguousarray(psi, dtype=np.float64), gl, gm) != 0:
            _err()
        return gl.reshape(self.E, self.Nq), gm.reshape(self.E, self.Nq)

    def inverse_mass(self, x):
        y = np.zeros(self.n)
        if lib().oracle_op_inverse_mass(self.h, np.ascontiguousarray(x, dtype=np.float64), y) != 0:
            _err()
        return y

    def wave_rhs(self, x, f_q=None):
        y = np.zeros(self.n)
        fq = None if f_q is None else np.ascontiguousarray(f_q, dtype=np.float64)
        if lib().oracle_op_wave_rhs(self.h, np.ascontiguousarray(x, dtype=np.float64), _ptr(fq), y) != 0:
            _err()
        return y

    def eval_volume(self, x):
        """Discrete solution at the volume points (E, Q, Nq) and the weights W|J| (E, Nq)."""
        Q = self.D if self.kind == 0 else 1
        out = np.zeros(self.E * Q * self.Nq)
        wj = np.zeros(self.E * self.Nq)
        if lib().oracle_op_eval_volume(self.h, np.ascontiguousarray(x, dtype=np.float64), out,
                                       wj.ctypes.data_as(C.c_void_p)) != 0:
            _err()
        return out.reshape(self.E, Q, self.Nq), wj.reshape(self.E, self.Nq)

    def facets(self):
        out = np.zeros(self.F * 5, dtype=np.int64)
        lib().oracle_op_facets(self.h, out)
        return out.reshape(self.F, 5)

    def array(self, name):
        n = lib().oracle_op_array_size(self.h, name.encode())
        out = np.zeros(n)
        if lib().oracle_op_array(self.h, name.encode(), out) != 0:
            _err()
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_op_free(self.h)
            self.h = None


def nodal_interpolation(D, to, frm):
    """I (n_to, n_from): the degree-frm nodal basis at the degree-to warp-and-blend nodes."""
    from math import comb
    out = np.zeros(comb(to + D, D) * comb(frm + D, D))
    if lib().oracle_nodal_interpolation(D, to, frm, out) != 0:
        _err()
    return out.reshape(comb(to + D, D), comb(frm + D, D))
