"""Shared test inputs: the same mesh, geometry nodes, coefficients and vectors for the HIP operator
and the CPU oracle."""
import numpy as np

import oracle_lib as O
import tandem_amd as T

BC_D, BC_N = 5, 1


def warp_nodes(nodes, amp=0.02):
    """Curvilinear map used throughout: x + amp sin(2 pi x) sin(2 pi y) [sin(2 pi z)] (1, ..., 1)."""
    s = amp * np.prod(np.sin(2.0 * np.pi * nodes), axis=-1)
    return nodes + s[..., None]


def smooth_coefficients(X, seed):
    """Smooth positive lambda/mu (or K) fields at points X (..., D)."""
    rng = np.random.default_rng(seed)
    D = X.shape[-1]
    k = rng.normal(size=(4, D)) * 3.0
    ph = rng.uniform(0, 2 * np.pi, size=4)
    g = np.cos(X @ k.T + ph) @ rng.uniform(0.1, 0.3, size=4)
    lam = 2.0 * np.exp(g)
    mu = 1.0 * np.exp(0.5 * np.sin(X @ k[0] + 1.0))
    return lam, mu


class Case:
    def __init__(self, D, n, scheme, N, kind, bc=BC_D, warp=0.02, seed=0):
        bcs = np.broadcast_to(np.asarray(bc, dtype=np.int32), (2 * D)).copy()
        self.pm = T.Mesh.box(D, n, scheme, bcs)
        self.om = O.Mesh.box(D, np.broadcast_to(n, (D,)), scheme, bcs)
        nodes = self.pm.affine_nodes(N)
        if warp:
            nodes = warp_nodes(nodes, warp)
        self.nodes = np.ascontiguousarray(nodes)
        self.op = T.DGOperator(self.pm, kind, N, self.nodes)
        X = self.op.volume_points()
        lam, mu = smooth_coefficients(X, seed)
        self.c0, self.c1 = (lam, mu) if kind == 0 else (lam, None)
        self.op.set_material(self.c0, self.c1)
        self.oracle = O.Operator(kind, self.om, N, self.c0, self.c1, nodes=self.nodes)
        self.n = self.op.number_of_local_dofs()
        assert self.n == self.oracle.n


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fault_plane_mesh(D, n, scheme, bc=BC_D, fault_bc=3, x0=0.5):
    """Box mesh (product and oracle) whose interior facets on the plane x = x0 carry fault_bc
    (tandem tags faults through the facet BC list, DGOperatorTopo.cpp:13-109)."""
    pm = T.Mesh.box(D, n, scheme, bc)
    v = pm.vertices
    faces = set()
    for el in pm.elements:
        for skip in range(D + 1):
            f = tuple(sorted(np.delete(el, skip)))
            if np.all(np.abs(v[list(f), 0] - x0) < 1e-12):
                faces.add(f)
    faces = np.array(sorted(faces), dtype=np.int64).reshape(-1, D)
    bf = np.concatenate([pm.boundary_facets, faces])
    bbc = np.concatenate([pm.boundary_bc, np.full(len(faces), fault_bc, dtype=np.int32)])
    return T.Mesh(v, pm.elements, bf, bbc), O.Mesh.from_arrays(v, pm.elements, bf, bbc)
