"""The allocation-free apply tdg_mult_noalloc (through the C ABI) against tdg_mult.

Two calls in a row on one GPU, the second on a smaller operator with natural faces elsewhere: the
shared scratch array still holds the first call's facet slots, which the second call must clear
where its faces carry no term.  Bar: bit-identical to tdg_mult (same kernels, same summation order).
An operator whose slots exceed the array is rejected with TDG_ERR_ARG, one not yet on the device
with TDG_ERR_STATE; nothing runs in either case.
"""
import ctypes as C

import numpy as np
import pytest

import tandem_amd as T
from tandem_amd._lib import lib
from tandem_amd import scenario as S

pytestmark = pytest.mark.gpu


def _operator(cells, bc, seed):
    mesh = T.Mesh.box(3, cells, scheme=1, bc=bc)
    nodes = S.warp(mesh.affine_nodes(4))
    op = T.DGOperator(mesh, T.ELASTICITY, 4, nodes)
    lam, mu = S.lame_fields(3, seed, length=1.0)
    X = op.volume_points()
    op.set_material(lam(X), mu(X))
    return op


def _noalloc(op, x, y):
    return lib().tdg_mult_noalloc(op._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()))


def test_noalloc_apply_equals_apply_on_two_calls(cuda):
    import torch

    ops = [_operator(4, [5, 1, 5, 5, 1, 5], 3), _operator(2, [1, 5, 5, 1, 5, 1], 4)]
    for i, op in enumerate(ops):
        assert lib().tdg_upload(op._h) == 0
        x = torch.from_numpy(np.random.default_rng(10 + i).uniform(-1.0, 1.0, op.n)).to(cuda)
        ref = torch.full_like(x, float("nan"))
        op.apply(x, ref)
        y = torch.full_like(x, float("nan"))
        assert _noalloc(op, x, y) == 0
        torch.cuda.synchronize()
        print("call %d: %d elements, max |diff| %.3e" % (i, op.E, float(torch.max(torch.abs(y - ref)))), flush=True)
        assert torch.equal(y, ref)


def test_noalloc_apply_rejects_what_it_cannot_hold(cuda):
    import torch

    op = _operator(2, 5, 5)
    x = torch.zeros(op.n, dtype=torch.float64, device=cuda)
    y = torch.zeros_like(x)
    assert _noalloc(op, x, y) == -2            # not on the device yet
    big = _operator(15, 5, 6)                  # 20,250 P4 tets: 8.5e6 slot entries > 2^23
    assert lib().tdg_upload(big._h) == 0
    xb = torch.zeros(big.n, dtype=torch.float64, device=cuda)
    yb = torch.full_like(xb, 7.0)
    assert _noalloc(big, xb, yb) == -1
    assert b"at most" in lib().tdg_last_error()
    torch.cuda.synchronize()
    assert torch.all(yb == 7.0)               # nothing ran
