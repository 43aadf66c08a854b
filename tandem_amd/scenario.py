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
