"""GPU parity of the symmetric half storage (sym=1): ffhip_mat_lower,
ffhip_spmv on a half matrix (HashMatrix.cpp:1128-1136), ffhip_mat_full and
ffhip_mat_set_half, against the oracle.

Bars: the half pattern is the lower triangle of the full pattern and its
values are the full matrix's, bit for bit (pure extraction); the half product
within 1e-14 of the oracle's half product (the device fuses multiply-adds);
the full matrix rebuilt from the half one equals the original bit for bit (the
device stiffness is symmetric bit for bit), so CG on it runs the same
iterations; an uploaded lower triangle declared half behaves alike; a matrix
with an entry above the diagonal cannot be declared half; CG refuses a half
matrix.
"""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _lower(rp, ci, va):
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci <= rows
    rl = np.zeros(n + 1, np.int64)
    np.add.at(rl, rows[keep] + 1, 1)
    return np.cumsum(rl), np.ascontiguousarray(ci[keep]), np.ascontiguousarray(va[keep])


@pytest.mark.parametrize("dims", [(5, 5, 5), (4, 7, 3)])
def test_lower_product_and_full(dims):
    import torch
    m = ffhip.cube(*dims)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.set_bc((1, 2, 3, 4, 5, 6))
    rp, ci, va = A.download()
    L = A.lower()
    assert L.half == 1 and A.half == 0
    rl, cl, vl = L.download()
    el, ecl, evl = _lower(rp, ci, va)
    assert np.array_equal(rl, el) and np.array_equal(cl, ecl) and np.array_equal(vl, evl)
    x = np.random.default_rng(1).standard_normal(A.n)
    xd = torch.tensor(x, device="cuda")
    yd = torch.zeros_like(xd)
    L.spmv(xd, yd)
    yo = ffo.spmv_half(rl, cl, vl, x)
    nb = np.abs(yo) < 1e20  # rows away from tgv
    assert np.abs(yd.cpu().numpy() - yo)[nb].max() <= 1e-14 * np.abs(yo[nb]).max() * 10
    yt = torch.zeros_like(xd)
    L.spmv(xd, yt, trans=True)
    assert torch.equal(yt, yd)
    F = L.full()
    rf, cf, vf = F.download()
    assert np.array_equal(rf, rp) and np.array_equal(cf, ci) and np.array_equal(vf, va)
    b = ffhip.rhs_p1(m, 1.0, (1, 2, 3, 4, 5, 6))
    x1, s1 = ffhip.CG(A).solve(b)
    x2, s2 = ffhip.CG(F).solve(b)
    assert s1["iters"] == s2["iters"] and np.array_equal(x1, x2)
    with pytest.raises(ffhip.FFHipError):
        ffhip.CG(L)


def test_uploaded_half_and_errors():
    import torch
    mo, rp, ci, va, onbc, b = ffo.poisson_p1_3d(4)
    rl, cl, vl = _lower(rp, ci, va)
    L = ffhip.Matrix.upload(rl, cl, vl)
    L.set_half(1)
    x = np.random.default_rng(2).standard_normal(rp.size - 1)
    xd = torch.tensor(x, device="cuda")
    yd = torch.zeros_like(xd)
    L.spmv(xd, yd)
    yo = ffo.spmv_half(rl, cl, vl, x)
    nb = np.abs(yo) < 1e20
    assert np.abs(yd.cpu().numpy() - yo)[nb].max() <= 1e-14 * np.abs(yo[nb]).max() * 10
    F = L.full()
    rf, cf, vf = F.download()
    assert np.array_equal(rf, rp) and np.array_equal(cf, ci)
    # the oracle's stiffness is symmetric only to rounding: the rebuilt upper
    # triangle mirrors the lower one
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    low = ci <= rows
    assert np.array_equal(vf[low], va[low])
    assert np.abs(vf - va).max() <= 1e-14 * np.abs(va[np.abs(va) < 1e20]).max()
    I, J, C = L.coo()
    assert np.array_equal(I, np.repeat(np.arange(rl.size - 1), np.diff(rl)))
    Afull = ffhip.Matrix.upload(rp, ci, va)
    with pytest.raises(ffhip.FFHipError):
        Afull.set_half(1)  # entries above the diagonal
