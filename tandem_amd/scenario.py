"""Scenario functionals for the benchmark and test problems (the role of tandem's Lua scenarios,
app/common/Scenario.h:19-119, ElasticityScenario.h): curvilinear warp, seeded smooth log-normal
Lame fields with analytic gradients, and manufactured solutions with their body forces.

All functions are vectorised numpy over points of shape (..., dim).
"""
import numpy as np


def warp(X, amplitude=0.02):
    """Curvilinear box map x -> x + a sin(2 pi x) sin(2 pi y) [sin(2 pi z)] (1, ..., 1).

    The displacement vanishes on the faces of the unit box, so the boundary planes are preserved;
    only interior geometry nodes move.  (The direction (1, ..., 1) is this project's choice; the
    benchmark statement fixes the amplitude and the sine product only.)
    """
    s = amplitude * np.prod(np.sin(2.0 * np.pi * X), axis=-1)
    return X + s[..., None]


class LogNormalField:
    """Smooth random field f(x) = mean * exp(sigma g(x) - sigma^2 / 2), E[f] = mean.

    g is a stationary unit-variance Gaussian field with covariance exp(-|r|^2 / (2 l^2)),
    approximated by M random Fourier features g(x) = sqrt(2/M) sum_m cos(w_m . x + b_m),
    w_m ~ N(0, l^-2 I), b_m ~ U(0, 2 pi), drawn from `seed`.
    """

    def __init__(self, dim, mean, sigma, length, seed, features=64):
        rng = np.random.default_rng(seed)
        self.w = rng.normal(scale=1.0 / length, size=(features, dim))
        self.b = rng.uniform(0.0, 2.0 * np.pi, size=features)
        self.mean, self.sigma = float(mean), float(sigma)
        self.scale = np.sqrt(2.0 / features)

    def _g(self, X, grad=False, chunk=1 << 18):
        shp = X.shape[:-1]
        P = X.reshape(-1, X.shape[-1])
        g = np.empty(P.shape[0])
        dg = np.empty_like(P) if grad else None
        for s in range(0, P.shape[0], chunk):
            arg = P[s:s + chunk] @ self.w.T + self.b
            g[s:s + chunk] = self.scale * np.cos(arg).sum(axis=1)
            if grad:
                dg[s:s + chunk] = -self.scale * (np.sin(arg) @ self.w)
        return g.reshape(shp), (dg.reshape(X.shape) if grad else None)

    def __call__(self, X):
        g, _ = self._g(X)
        return self.mean * np.exp(self.sigma * g - 0.5 * self.sigma ** 2)

    def value_and_grad(self, X):
        g, dg = self._g(X, grad=True)
        f = self.mean * np.exp(self.sigma * g - 0.5 * self.sigma ** 2)
        return f, (f * self.sigma)[..., None] * dg


def lame_fields(dim, seed, length=0.2, sigma=0.3):
    """The benchmark's lambda (mean 2) and mu (mean 1) fields."""
    return (LogNormalField(dim, 2.0, sigma, length, seed), LogNormalField(dim, 1.0, sigma, length, seed + 1))


# ------------------------------------------------------------------ manufactured solutions
def cosine_solution(X):
    """u = (cos(pi x) cos(pi y), 0, cos(pi x) cos(pi z)) (examples/elasticity/3d/cosine.lua:33-38)
    in 3D; (cos(pi x) cos(pi y), cos(pi x) sin(pi y)) in 2D.  Returns u (..., D) and
    grad u (..., D, D) with grad[..., i, j] = d u_i / d x_j."""
    pi = np.pi
    D = X.shape[-1]
    x, y = X[..., 0], X[..., 1]
    u = np.zeros(X.shape)
    G = np.zeros(X.shape + (D,))
    if D == 3:
        z = X[..., 2]
        u[..., 0] = np.cos(pi * x) * np.cos(pi * y)
        u[..., 2] = np.cos(pi * x) * np.cos(pi * z)
        G[..., 0, 0] = -pi * np.sin(pi * x) * np.cos(pi * y)
        G[..., 0, 1] = -pi * np.cos(pi * x) * np.sin(pi * y)
        G[..., 2, 0] = -pi * np.sin(pi * x) * np.cos(pi * z)
        G[..., 2, 2] = -pi * np.cos(pi * x) * np.sin(pi * z)
    else:
        u[..., 0] = np.cos(pi * x) * np.cos(pi * y)
        u[..., 1] = np.cos(pi * x) * np.sin(pi * y)
        G[..., 0, 0] = -pi * np.sin(pi * x) * np.cos(pi * y)
        G[..., 0, 1] = -pi * np.cos(pi * x) * np.sin(pi * y)
        G[..., 1, 0] = -pi * np.sin(pi * x) * np.sin(pi * y)
        G[..., 1, 1] = pi * np.cos(pi * x) * np.cos(pi * y)
    return u, G


def _cosine_second(X):
    """Second derivatives H[..., i, j, k] = d^2 u_i / dx_j dx_k of cosine_solution."""
    pi2 = np.pi ** 2
    D = X.shape[-1]
    x, y = np.pi * X[..., 0], np.pi * X[..., 1]
    H = np.zeros(X.shape + (D, D))
    if D == 3:
        z = np.pi * X[..., 2]
        H[..., 0, 0, 0] = -pi2 * np.cos(x) * np.cos(y)
        H[..., 0, 1, 1] = -pi2 * np.cos(x) * np.cos(y)
        H[..., 0, 0, 1] = H[..., 0, 1, 0] = pi2 * np.sin(x) * np.sin(y)
        H[..., 2, 0, 0] = -pi2 * np.cos(x) * np.cos(z)
        H[..., 2, 2, 2] = -pi2 * np.cos(x) * np.cos(z)
        H[..., 2, 0, 2] = H[..., 2, 2, 0] = pi2 * np.sin(x) * np.sin(z)
    else:
        H[..., 0, 0, 0] = -pi2 * np.cos(x) * np.cos(y)
        H[..., 0, 1, 1] = -pi2 * np.cos(x) * np.cos(y)
        H[..., 0, 0, 1] = H[..., 0, 1, 0] = pi2 * np.sin(x) * np.sin(y)
        H[..., 1, 0, 0] = -pi2 * np.cos(x) * np.sin(y)
        H[..., 1, 1, 1] = -pi2 * np.cos(x) * np.sin(y)
        H[..., 1, 0, 1] = H[..., 1, 1, 0] = -pi2 * np.sin(x) * np.cos(y)
    return H


def elasticity_body_force(X, lam, mu):
    """f = -div sigma(u*) for the cosine solution with variable Lame fields.

    lam, mu: callables with value_and_grad(X) (LogNormalField) or constants.
    sigma = lam tr(eps) I + 2 mu eps, eps = (grad u + grad u^T) / 2.
    Returns f with shape (..., D).
    """
    D = X.shape[-1]
    _, G = cosine_solution(X)
    H = _cosine_second(X)

    def vg(c):
        if hasattr(c, "value_and_grad"):
            return c.value_and_grad(X)
        return np.full(X.shape[:-1], float(c)), np.zeros(X.shape)

    l, dl = vg(lam)
    m, dm = vg(mu)
    div = np.trace(G, axis1=-2, axis2=-1)
    grad_div = np.einsum("...iij->...j", H)
    lap = np.einsum("...ijj->...i", H)
    S = G + np.swapaxes(G, -1, -2)  # 2 eps
    # div sigma_i = d_i(lam div) + d_j(mu S_ij)
    ds = dl * div[..., None] + l[..., None] * grad_div + np.einsum("...j,...ij->...i", dm, S) + m[..., None] * (lap + grad_div)
    return -ds


def poisson_sine(X):
    """u = prod sin(pi x_d), f = D pi^2 u (K = 1); zero Dirichlet data on the unit box."""
    D = X.shape[-1]
    u = np.prod(np.sin(np.pi * X), axis=-1)
    return u, D * np.pi ** 2 * u


class FaultRectangle:
    """2D plane-strain Green's-function problem: the rectangle [0, W] x [-H, 0] with a planar fault
    dipping at `dip` degrees from the surface point (x0, 0), made of `nfault` straight facets.

    Triangulation: a seeded, jittered point cloud of spacing h (exact points on the rectangle's
    sides and on the fault, none closer than `band` h to the fault) triangulated by Delaunay; the
    fault segments are then edges of the triangulation (checked).  BCs: the surface y = 0 is free
    (natural), the other sides are Dirichlet, the fault facets are Fault.  (Mesh and BC choices are
    this project's; the reference reads such meshes from gmsh.)
    """

    def __init__(self, W=400.0, H=200.0, h=0.9, nfault=512, fault_h=0.5, dip=30.0, x0=80.0, seed=20261205,
                 band=0.6):
        from scipy.spatial import Delaunay

        rng = np.random.default_rng(seed)
        d = np.array([np.cos(np.radians(dip)), -np.sin(np.radians(dip))])
        fault = np.array([x0, 0.0])[None, :] + fault_h * np.arange(nfault + 1)[:, None] * d[None, :]
        if fault[-1, 1] <= -H or fault[-1, 0] >= W:
            raise ValueError("fault leaves the rectangle")
        nx, ny = int(round(W / h)), int(round(H / h))
        xs, ys = np.linspace(0.0, W, nx + 1), np.linspace(-H, 0.0, ny + 1)
        side = [np.stack([xs, np.full_like(xs, -H)], 1), np.stack([xs, np.zeros_like(xs)], 1),
                np.stack([np.zeros_like(ys), ys], 1), np.stack([np.full_like(ys, W), ys], 1)]
        gx, gy = np.meshgrid(xs[1:-1], ys[1:-1], indexing="ij")
        inner = np.stack([gx.ravel(), gy.ravel()], 1) + rng.uniform(-0.3 * h, 0.3 * h, (gx.size, 2))
        # drop points close to the fault segment (distance to the segment, not the line)
        rel = inner - fault[0]
        t = np.clip(rel @ d, 0.0, fault_h * nfault)
        dist = np.linalg.norm(rel - t[:, None] * d[None, :], axis=1)
        inner = inner[dist > band * h]
        surf = side[1]
        surf = surf[np.abs(surf[:, 0] - x0) > 0.5 * h]  # the fault's surface point replaces its neighbour
        pts = np.concatenate([side[0], surf, side[2], side[3], inner, fault])
        pts = np.unique(np.round(pts, 12), axis=0)
        tri = Delaunay(pts)
        elems = np.sort(tri.simplices.astype(np.int64), axis=1)
        # keep non-degenerate triangles
        p = pts[elems]
        area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                            (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
        elems = elems[area > 1e-12 * h * h]
        index = {tuple(np.round(q, 12)): i for i, q in enumerate(pts)}
        fv = np.array([index[tuple(np.round(q, 12))] for q in fault])
        fedges = np.sort(np.stack([fv[:-1], fv[1:]], 1), axis=1)
        edges = np.sort(np.concatenate([elems[:, [0, 1]], elems[:, [0, 2]], elems[:, [1, 2]]]), axis=1)
        eset = set(map(tuple, edges))
        missing = [tuple(e) for e in fedges if tuple(e) not in eset]
        if missing:
            raise RuntimeError("%d fault segments are not triangulation edges" % len(missing))
        # boundary edges (appear once) with their BC
        uniq, cnt = np.unique(edges, axis=0, return_counts=True)
        bnd = uniq[cnt == 1]
        mid = pts[bnd].mean(axis=1)
        bc = np.where(np.abs(mid[:, 1]) < 1e-9, 1, 5).astype(np.int32)
        self.vertices, self.elements = pts, elems
        self.boundary_facets = np.concatenate([bnd, fedges])
        self.boundary_bc = np.concatenate([bc, np.full(len(fedges), 3, dtype=np.int32)])
        self.fault_vertices = fv
        self.tangent = d

    def mesh(self):
        from .operator import Mesh
        return Mesh(self.vertices, self.elements, self.boundary_facets, self.boundary_bc)

    def slip_basis(self, op, columns=None):
        """Facet data (ncols, F, 2, nq) of the slip basis functions: facet-nodal Lagrange polynomials
        of degree N on each fault facet (nodes at the Gauss-Lobatto points of the facet), times the
        along-fault unit tangent, oriented so that the slip is the hanging wall's displacement
        minus the footwall's ([[u]] = u(side 0) - u(side 1), hence the sign by side).
        Column c = (fault facet c // (N+1), node c % (N+1)); `columns` selects a subset."""
        N = op.degree
        if not hasattr(self, "_slip"):
            fct = op.facets()                    # (F, 5): elem0, elem1, local0, local1, bc
            Xf = op.facet_points()               # (F, nq, 2)
            faults = np.nonzero(fct[:, 4] == 3)[0]
            # Gauss-Lobatto points on [0, 1]: the end points and the roots of P_N'
            lg = np.polynomial.legendre.Legendre.basis(N).deriv().roots()
            nodes1d = np.sort(np.concatenate([[0.0, 1.0], 0.5 * (np.real(lg) + 1.0)]))
            # each fault facet's end points (its two mesh vertices), matched by midpoint
            fv = self.vertices[self.boundary_facets[self.boundary_bc == 3]]      # (nf, 2, 2)
            mids = fv.mean(axis=1)
            fmid = Xf[faults].mean(axis=1)
            match = np.argmin(np.linalg.norm(fmid[:, None, :] - mids[None, :, :], axis=2), axis=1)
            seg = fv[match]                                                        # (nf, 2, 2)
            dv = seg[:, 1] - seg[:, 0]
            t = np.einsum("fqd,fd->fq", Xf[faults] - seg[:, None, 0], dv) / np.einsum("fd,fd->f", dv, dv)[:, None]
            L = np.ones((len(faults), N + 1, t.shape[1]))                          # l_i(t_q)
            for i in range(N + 1):
                for m in range(N + 1):
                    if m != i:
                        L[:, i] *= (t - nodes1d[m]) / (nodes1d[i] - nodes1d[m])
            cent = self.vertices[self.elements].mean(axis=1)
            nrm = np.array([-self.tangent[1], self.tangent[0]])                    # towards the hanging wall
            up = np.einsum("fd,d->f", cent[fct[faults, 0]] - fmid, nrm) > 0
            self._slip = (faults, L, np.where(up, 1.0, -1.0))
        faults, L, sgn = self._slip
        ncols_all = len(faults) * (N + 1)
        cols = np.arange(ncols_all) if columns is None else np.asarray(columns)
        out = np.zeros((len(cols), op.F, 2, op.nq))
        fi, ni = cols // (N + 1), cols % (N + 1)
        out[np.arange(len(cols)), faults[fi]] = (sgn[fi][:, None, None] * L[fi, ni][:, None, :] *
                                                 self.tangent[None, :, None])
        return out

    def uniform_slip(self, op):
        """Facet data (F, 2, nq) of a unit along-fault slip on every fault facet (the sum of all slip
        basis functions), with the sign convention of slip_basis."""
        self.slip_basis(op, [0])
        faults, _, sgn = self._slip
        out = np.zeros((op.F, 2, op.nq))
        out[faults] = sgn[:, None, None] * self.tangent[None, :, None]
        return out

    @staticmethod
    def with_elements(n_elements, seed):
        """A fault rectangle of the default geometry meshed with about n_elements triangles (spacing
        from 2 W H / h^2 triangles; fault facets of about 0.55 h along the same 256-long fault)."""
        h = float(np.sqrt(2.0 * 400.0 * 200.0 / n_elements))
        fh = 0.55 * h
        return FaultRectangle(h=h, nfault=max(4, int(round(256.0 / fh))), fault_h=256.0 / max(4, int(round(256.0 / fh))),
                              seed=seed)

