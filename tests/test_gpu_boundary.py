"""GPU parity of the P1 boundary integrals (ffhip_assemble_p1_boundary,
ffhip_rhs_p1_boundary: int2d(Th3, labels)(alpha*u*v) and (g*v), int1d in 2D)
against the oracle, which follows FreeFEM's face loop (problem.cpp:1273-1327,
6515-6561, 8517-8590).

Bars: matrix values within 1e-12 of the largest non-tgv entry (tgv entries
exact) after volume term + boundary term + on(...); right-hand sides within
1e-12 relative; the Robin problem solved to eps = 1e-10 within 1e-8 of the
oracle's solution.  The GPU sums each boundary element's rule in its own
vertex order, the oracle in the element's face order: only rounding differs.
"""
import numpy as np
import pytest

import ffo
from test_gpu_p1 import _vals_close

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _own_points(mo, dim):
    be = (mo["btri"] if dim == 3 else mo["bedges"]).reshape(-1, dim)
    X = (mo["xyz"] if dim == 3 else mo["xy"]).reshape(-1, dim)
    if dim == 3:
        qp, qw = np.zeros(14), np.zeros(7)
        ffo.lib().ffo_qf_tri5(qp, qw)
        lam = np.stack([1 - qp[0::2] - qp[1::2], qp[0::2], qp[1::2]], 1)
    else:
        x, w = np.zeros(3), np.zeros(3)
        ffo.lib().ffo_qf_gl3(x, w)
        lam = np.stack([1 - x, x], 1)
    return np.einsum("qk,fkc->fqc", lam, X[be])


def _coef(case, mo, dim):
    rng = np.random.default_rng(11)
    if case == "const":
        return ffo.COEF_CONST, 2.5, None
    if case == "sin3":
        return ffo.COEF_SIN3, 0.0, None
    if case == "p0":
        return ffo.COEF_P0, 0.0, rng.uniform(0.5, 2.0, mo["nbe"])
    P = _own_points(mo, dim)
    return ffo.COEF_QP, 0.0, np.ascontiguousarray(1.0 + P[:, :, 0] ** 2 + 0.5 * P[:, :, 1])


CASES = ["const", "sin3", "p0", "qp"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dim", [3, 2])
def test_robin_matrix_vs_oracle(case, dim):
    if dim == 3:
        mo, m = ffo.cube(6), ffhip.cube(6)
        el, nloc, robin, dirichlet = mo["tets"], 4, (1, 2, 3), (5, 6)
    else:
        mo, m = ffo.square(9, 7), ffhip.square(9, 7)
        el, nloc, robin, dirichlet = mo["tris"], 3, (1, 2), (4,)
    kind, c0, ce = _coef(case, mo, dim)
    rp, ci = ffo.pattern(mo["nv"], el, nloc)
    if dim == 3:
        vo = ffo.assemble_p1_3d(mo, rp, ci, ffo.COEF_SIN3)
    else:
        vo = ffo.assemble_p1_2d(mo, rp, ci, ffo.COEF_SIN3)
    ffo.p1_boundary(mo, robin, kind, c0, coef=ce, rowptr=rp, colidx=ci, vals=vo)
    be = mo["btri"] if dim == 3 else mo["bedges"]
    ffo.set_bc(rp, ci, vo, ffo.mark_bc(mo["nv"], be, mo["blab"], dim, dirichlet))
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.assemble_p1_boundary(robin, kind, c0, ce)
    A.set_bc(dirichlet)
    rpa, cia, va = A.download()
    assert np.array_equal(rpa, rp) and np.array_equal(cia, ci)
    assert _vals_close(va, vo)
    # the boundary part alone is not negligible next to the stiffness
    B = ffhip.Matrix.p1(m)
    B.assemble_p1(ffhip.COEF_SIN3)
    _, _, vb = B.download()
    assert np.abs(va - vb)[vo < 1e29].max() > 1e-3 * np.abs(vb).max()


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dim", [3, 2])
def test_neumann_rhs_vs_oracle(case, dim):
    if dim == 3:
        mo, m = ffo.cube(5, 4, 6), ffhip.cube(5, 4, 6)
        neumann, dirichlet = (2, 4, 6), (1, 3)
        bo = ffo.rhs_p1_3d(mo, 1.0)
    else:
        mo, m = ffo.square(8, 11), ffhip.square(8, 11)
        neumann, dirichlet = (1, 3), (2,)
        bo = ffo.rhs_p1_2d(mo, 1.0)
    kind, g0, ce = _coef(case, mo, dim)
    ffo.p1_boundary(mo, neumann, kind, g0, coef=ce, b=bo)
    be = mo["btri"] if dim == 3 else mo["bedges"]
    ffo.rhs_bc(bo, ffo.mark_bc(mo["nv"], be, mo["blab"], dim, dirichlet), 1e30, 0.25)
    b = ffhip.rhs_p1(m, 1.0, dirichlet, 1e30, 0.25)
    ffhip.rhs_p1_boundary(m, b, neumann, kind, g0, ce, dirichlet=dirichlet)
    big = bo >= 1e29
    assert np.array_equal(b[big], bo[big]) and big.any()
    assert np.abs(b - bo)[~big].max() <= 1e-12 * np.abs(bo[~big]).max()


def test_device_arrays_and_robin_solve():
    import torch
    mo, m = ffo.cube(8), ffhip.cube(8)
    robin, neumann, dirichlet = (1, 2), (3, 4), (5, 6)
    kind, _, ce = _coef("qp", mo, 3)
    rp, ci = ffo.pattern(mo["nv"], mo["tets"], 4)
    vo = ffo.assemble_p1_3d(mo, rp, ci, ffo.COEF_SIN3)
    ffo.p1_boundary(mo, robin, kind, 0.0, coef=ce, rowptr=rp, colidx=ci, vals=vo)
    onbc = ffo.mark_bc(mo["nv"], mo["btri"], mo["blab"], 3, dirichlet)
    ffo.set_bc(rp, ci, vo, onbc)
    bo = ffo.rhs_p1_3d(mo, 1.0)
    ffo.p1_boundary(mo, neumann, ffo.COEF_SIN3, 0.0, b=bo)
    ffo.rhs_bc(bo, onbc, 1e30, 0.0)
    xo, so = ffo.pcg(rp, ci, vo, bo, eps=1e-10)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3)
    A.assemble_p1_boundary(robin, kind, 0.0, torch.tensor(ce.ravel(), device="cuda"))
    A.set_bc(dirichlet)
    bd = torch.zeros(mo["nv"], dtype=torch.float64, device="cuda")
    ffhip.rhs_p1(m, 1.0, dirichlet, 1e30, 0.0, out=bd)
    ffhip.rhs_p1_boundary(m, bd, neumann, ffhip.COEF_SIN3, 0.0, dirichlet=dirichlet)
    b = bd.cpu().numpy()
    assert np.abs(b - bo)[bo < 1e29].max() <= 1e-12 * np.abs(bo[bo < 1e29]).max()
    x, st = ffhip.CG(A).solve(b, eps=1e-10)
    assert abs(st["iters"] - so["iters"]) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_unlisted_labels_and_errors():
    m = ffhip.cube(3)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_CONST, 1.0)
    _, _, v0 = A.download()
    A.assemble_p1_boundary((42,), ffhip.COEF_CONST, 1.0)
    _, _, v1 = A.download()
    assert np.array_equal(v0, v1)
    with pytest.raises(ffhip.FFHipError):
        A.assemble_p1_boundary((1,), ffhip.COEF_P0, 0.0, None)  # P0 without values
    # a boundary triangle that is no face of the mesh is reported
    mo = ffo.cube(3)
    bt = mo["btri"].copy()
    bt[0:3] = [0, 5, 63]
    M = ffhip.mesh3_upload(mo["xyz"], mo["tets"], bt, mo["blab"])
    B = ffhip.Matrix.p1(M)
    B.assemble_p1(ffhip.COEF_CONST, 1.0)
    with pytest.raises(ffhip.FFHipError):
        B.assemble_p1_boundary((int(mo["blab"][0]),), ffhip.COEF_CONST, 1.0)
