"""Adjoint gradient kernel (through the C ABI) against the oracle's material gradient.

The oracle's gradient (d/d(lam_Q, mu_Q) of psi^T A u) is pinned by Richardson-extrapolated central
differences of the oracle's own apply in the CPU tests.  Bar: FP64 relative L2 <= 1e-11; FP32 <= 1e-4.
"""
import numpy as np
import pytest

from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

CASES = [(3, 2, 1, 4, [5] * 6), (3, 2, 0, 2, [5, 1, 1, 5, 3, 1]), (2, 3, 0, 6, [5, 1, 3, 5]), (2, 4, 0, 2, [5] * 4)]


@pytest.mark.parametrize("case", CASES, ids=["D%d-n%d-s%d-P%d" % c[:4] for c in CASES])
def test_material_gradient_matches_oracle(cuda, case):
    import torch

    D, n, scheme, N, bc = case
    c = Case(D, n, scheme, N, 0, bc=bc, seed=9)
    rng = np.random.default_rng(12)
    u = rng.uniform(-1, 1, c.n)
    psi = rng.uniform(-1, 1, c.n)
    rl, rm = c.oracle.material_gradient(u, psi)
    print("case", case, flush=True)
    gl, gm = c.op.material_gradient(torch.from_numpy(u).to(cuda), torch.from_numpy(psi).to(cuda))
    el, em = rel_l2(gl.cpu().numpy(), rl), rel_l2(gm.cpu().numpy(), rm)
    print("  fp64 rel L2 lam %.3e mu %.3e" % (el, em), flush=True)
    assert el <= 1e-11 and em <= 1e-11
    gl, gm = c.op.material_gradient(torch.from_numpy(u).float().to(cuda), torch.from_numpy(psi).float().to(cuda))
    el, em = rel_l2(gl.double().cpu().numpy(), rl), rel_l2(gm.double().cpu().numpy(), rm)
    print("  fp32 rel L2 lam %.3e mu %.3e" % (el, em), flush=True)
    assert el <= 1e-4 and em <= 1e-4
