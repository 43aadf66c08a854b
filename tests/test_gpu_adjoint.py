"""Adjoint gradient kernel (through the C ABI) against the oracle's material gradient.

The oracle's gradient (d/d(lam_Q, mu_Q) of psi^T A u) is pinned by Richardson-extrapolated central
differences of the oracle's own apply in the CPU tests.  Bar: FP64 relative L2 <= 1e-11; FP32 <= 1e-4.  Both kernels (matrix-core and thread-per-point) are
checked; the element counts are not multiples of the 16 elements of one wave.
"""
import numpy as np
import pytest

from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

CASES = [(3, 3, 1, 4, [5] * 6), (3, 2, 0, 2, [5, 1, 1, 5, 3, 1]), (2, 3, 0, 6, [5, 1, 3, 5]), (2, 5, 0, 2, [5] * 4), (2, 4, 0, 3, [5, 1, 5, 5])]


@pytest.mark.parametrize("variant", ["auto", "mfma-xcd", "points"])
@pytest.mark.parametrize("case", CASES, ids=["D%d-n%d-s%d-P%d" % c[:4] for c in CASES])
def test_material_gradient_matches_oracle(cuda, case, variant):
    """auto and mfma-xcd run the matrix-core gradient kernel (every case here has one), points the
    thread-per-point kernel."""
    import torch

    D, n, scheme, N, bc = case
    c = Case(D, n, scheme, N, 0, bc=bc, seed=9)
    c.op.set_variant(variant)
    rng = np.random.default_rng(12)
    u = rng.uniform(-1, 1, c.n)
    psi = rng.uniform(-1, 1, c.n)
    rl, rm = c.oracle.material_gradient(u, psi)
    print("case", case, variant, "elements", c.op.E, flush=True)
    gl, gm = c.op.material_gradient(torch.from_numpy(u).to(cuda), torch.from_numpy(psi).to(cuda))
    el, em = rel_l2(gl.cpu().numpy(), rl), rel_l2(gm.cpu().numpy(), rm)
    print("  fp64 rel L2 lam %.3e mu %.3e" % (el, em), flush=True)
    assert el <= 1e-11 and em <= 1e-11
    gl, gm = c.op.material_gradient(torch.from_numpy(u).float().to(cuda), torch.from_numpy(psi).float().to(cuda))
    el, em = rel_l2(gl.double().cpu().numpy(), rl), rel_l2(gm.double().cpu().numpy(), rm)
    print("  fp32 rel L2 lam %.3e mu %.3e" % (el, em), flush=True)
    assert el <= 1e-4 and em <= 1e-4


def _smooth_direction(X, rng):
    """A seeded smooth relative perturbation field in [-1, 1] at the points X."""
    k = rng.normal(size=(3, X.shape[-1])) * 2.0
    ph = rng.uniform(0, 2 * np.pi, size=3)
    return np.tanh(np.cos(X @ k.T + ph) @ rng.uniform(0.3, 1.0, size=3))


@pytest.mark.parametrize("cells,N", [(1, 4), (2, 2)])
def test_misfit_gradient_matches_finite_differences(cuda, cells, N):
    """The adjoint gradient of J(lam, mu) = 1/2 |B u - d|^2, u = A(lam, mu)^-1 f, on a curved tet
    mesh with a free surface (z = 1): forward and adjoint by dense solves of the oracle's assembled
    matrix, gradient -(g_lam, g_mu)(u, psi) from the GPU kernel, against Richardson-extrapolated
    central differences of J (dense solves of re-assembled matrices) along 16 seeded smooth
    directions.  Bar: relative agreement 1e-6 in FP64."""
    import torch
    import oracle_lib as O

    c = Case(3, cells, 1, N, 0, bc=[5, 5, 5, 5, 5, 1], seed=11)
    X = c.op.volume_points()
    lam, mu = c.c0, c.c1
    rng = np.random.default_rng(40 + cells)
    f = rng.uniform(-1, 1, c.n)
    Bm = rng.normal(size=(64, c.n))
    d = rng.normal(size=64) * 0.01

    def solve(l, m, rhs):
        A = O.Operator(0, c.om, N, l, m, nodes=c.nodes).assemble()
        return np.linalg.solve(A, rhs)

    def J(l, m):
        r = Bm @ solve(l, m, f) - d
        return 0.5 * float(r @ r)

    u = solve(lam, mu, f)
    psi = solve(lam, mu, Bm.T @ (Bm @ u - d))
    gl, gm = c.op.material_gradient(torch.from_numpy(u).to(cuda), torch.from_numpy(psi).to(cuda))
    gl, gm = -gl.cpu().numpy(), -gm.cpu().numpy()
    worst = 0.0
    for k in range(16):
        dl = lam * _smooth_direction(X, rng)
        dm = mu * _smooth_direction(X, rng)
        an = float(np.sum(gl * dl) + np.sum(gm * dm))
        h = 1e-4  # small enough that no penalty extremal node changes, large against rounding
        c1 = (J(lam + h * dl, mu + h * dm) - J(lam - h * dl, mu - h * dm)) / (2 * h)
        c2 = (J(lam + h / 2 * dl, mu + h / 2 * dm) - J(lam - h / 2 * dl, mu - h / 2 * dm)) / h
        fd = (4 * c2 - c1) / 3
        worst = max(worst, abs(fd - an) / abs(an))
    print("cells", cells, "P%d" % N, "n", c.n, "max rel err %.3e" % worst, flush=True)
    assert worst <= 1e-6
