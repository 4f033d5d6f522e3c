"""GPU parity of the P1 form with a reaction term,
int3d(Th)(kappa*grad(u)'*grad(v) + m*u*v) (ffhip_assemble_p1_form), against
the oracle: values within 1e-12 of the largest entry (tgv entries exact), in
3D (constant, analytic, P0 and quadrature-point coefficients; the half-wave
and the lane-per-row gathers) and in 2D; an implicit heat step solved to
eps = 1e-10 within 1e-8 of the oracle's solution."""
import numpy as np
import pytest

import ffo
from test_gpu_p1 import _fan_mesh, _vals_close

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")
LABELS = (1, 2, 3, 4, 5, 6)


@pytest.mark.parametrize("case", ["const", "sin3", "p0"])
def test_form_values_vs_oracle(case):
    n = 6
    mo = ffo.cube(n)
    rp, ci = ffo.pattern(mo["nv"], mo["tets"], 4)
    me = np.exp(np.random.default_rng(4).standard_normal(mo["nt"])) if case == "p0" else None
    mkind = {"const": ffo.COEF_CONST, "sin3": ffo.COEF_SIN3, "p0": ffo.COEF_P0}[case]
    vo = ffo.assemble_p1_3d(mo, rp, ci, ffo.COEF_SIN3, mkind=mkind, m0=2.5, mass_e=me)
    m = ffhip.cube(n)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3, 1.0, None, mkind, 2.5, me)
    rpa, cia, va = A.download()
    assert np.array_equal(rpa, rp) and np.array_equal(cia, ci)
    assert _vals_close(va, vo)


def test_form_quadrature_point_coefficient_and_device_arrays():
    import torch
    n = 5
    m = ffhip.cube(n)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_CONST, 1.0, None, ffhip.COEF_SIN3, 0.0, None)
    _, _, v_an = A.download()
    # the same coefficient evaluated at the 14 Tet_5 points of every element
    mo = ffo.cube(n)
    pts, w = np.zeros(42), np.zeros(14)
    ffo.lib().ffo_qf_tet5(pts, w)
    X = mo["xyz"].reshape(-1, 3)[mo["tets"].reshape(-1, 4)]  # nt x 4 x 3
    P = pts.reshape(14, 3)
    lam = np.c_[1 - (P[:, 0] + P[:, 1] + P[:, 2]), P]  # 14 x 4
    Q = np.einsum("qi,eik->eqk", lam, X)
    mq = 1 + 0.5 * np.sin(2 * np.pi * Q[..., 0]) * np.sin(2 * np.pi * Q[..., 1]) * np.sin(2 * np.pi * Q[..., 2])
    B = ffhip.Matrix.p1(m)
    B.assemble_p1(ffhip.COEF_CONST, 1.0, None, ffhip.COEF_QP, 0.0,
                  torch.tensor(mq.ravel(), device="cuda"))
    _, _, v_qp = B.download()
    assert np.abs(v_qp - v_an).max() <= 1e-13 * np.abs(v_an).max()


def test_form_long_rows_and_2d():
    fm = _fan_mesh(20)
    rp, ci = ffo.pattern(fm["nv"], fm["tets"], 4)
    vo = ffo.assemble_p1_3d(fm, rp, ci, ffo.COEF_CONST, 1.0, mkind=ffo.COEF_CONST, m0=3.0)
    M = ffhip.mesh3_upload(fm["xyz"], fm["tets"], fm["btri"], fm["blab"])
    A = ffhip.Matrix.p1(M)
    A.assemble_p1(ffhip.COEF_CONST, 1.0, None, ffhip.COEF_CONST, 3.0, None)
    _, _, va = A.download()
    assert np.abs(va - vo).max() <= 1e-12 * np.abs(vo).max()
    sq = ffo.square(20, 17)
    rp, ci = ffo.pattern(sq["nv"], sq["tris"], 3)
    vo = ffo.assemble_p1_2d(sq, rp, ci, ffo.COEF_SIN3, mkind=ffo.COEF_CONST, m0=0.5)
    m2 = ffhip.square(20, 17)
    A2 = ffhip.Matrix.p1(m2)
    A2.assemble_p1(ffhip.COEF_SIN3, 1.0, None, ffhip.COEF_CONST, 0.5, None)
    _, _, va = A2.download()
    assert np.abs(va - vo).max() <= 1e-12 * np.abs(vo).max()


def test_implicit_heat_step_solve():
    # (M/dt + K) u = rhs with Dirichlet tgv rows, as a time step of u_t - div(kappa grad u) = 1
    n, dt = 8, 0.01
    mo = ffo.cube(n)
    rp, ci = ffo.pattern(mo["nv"], mo["tets"], 4)
    vo = ffo.assemble_p1_3d(mo, rp, ci, ffo.COEF_SIN3, mkind=ffo.COEF_CONST, m0=1.0 / dt)
    onbc = ffo.mark_bc(mo["nv"], mo["btri"], mo["blab"], 3, LABELS)
    ffo.set_bc(rp, ci, vo, onbc)
    b = ffo.rhs_p1_3d(mo, 1.0)
    ffo.rhs_bc(b, onbc)
    xo, so = ffo.pcg(rp, ci, vo, b, eps=1e-10)
    m = ffhip.cube(n)
    A = ffhip.Matrix.p1(m)
    A.assemble_p1(ffhip.COEF_SIN3, 1.0, None, ffhip.COEF_CONST, 1.0 / dt, None)
    A.set_bc(LABELS)
    x, st = ffhip.CG(A).solve(b, eps=1e-10)
    assert st["converged"] == 1 and abs(st["iters"] - so["iters"]) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
