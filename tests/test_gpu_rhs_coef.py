"""GPU parity of the linear form with a coefficient and per-dof Dirichlet
values (ffhip_rhs_p1_form: Element_rhs problem.cpp:7940-7978 with f at the
quadrature points, AssembleBC's B rows problem.cpp:10120-10176) against the
oracle: within 1e-12 relative on the non-Dirichlet rows, the Dirichlet rows
tgv * g exactly, for constant, analytic, P0 and quadrature-point f in 3D and
2D, host and device arrays."""
import numpy as np
import pytest

import ffo

pytestmark = pytest.mark.gpu

ffhip = pytest.importorskip("ffhip")


def _qp(mo, dim):
    X = (mo["xyz"] if dim == 3 else mo["xy"]).reshape(-1, dim)
    el = (mo["tets"] if dim == 3 else mo["tris"]).reshape(-1, dim + 1)
    if dim == 3:
        qp, qw = np.zeros(42), np.zeros(14)
        ffo.lib().ffo_qf_tet5(qp, qw)
        Q = qp.reshape(14, 3)
        lam = np.c_[1 - (Q[:, 0] + Q[:, 1] + Q[:, 2]), Q]
    else:
        qp, qw = np.zeros(14), np.zeros(7)
        ffo.lib().ffo_qf_tri5(qp, qw)
        Q = qp.reshape(7, 2)
        lam = np.c_[1 - Q[:, 0] - Q[:, 1], Q]
    P = np.einsum("qi,eik->eqk", lam, X[el])
    return np.ascontiguousarray(np.exp(P[..., 0]) + P[..., 1] ** 2)


@pytest.mark.parametrize("case", ["const", "sin3", "p0", "qp"])
@pytest.mark.parametrize("dim", [3, 2])
def test_rhs_form_vs_oracle(case, dim):
    if dim == 3:
        mo, m, labels = ffo.cube(5, 6, 4), ffhip.cube(5, 6, 4), (1, 6)
    else:
        mo, m, labels = ffo.square(9, 8), ffhip.square(9, 8), (3,)
    kind = {"const": ffo.COEF_CONST, "sin3": ffo.COEF_SIN3, "p0": ffo.COEF_P0, "qp": ffo.COEF_QP}[case]
    fc = None
    if case == "p0":
        fc = np.random.default_rng(1).uniform(-1, 2, mo["nt"])
    elif case == "qp":
        fc = _qp(mo, dim)
    bo = ffo.rhs_p1_coef(mo, kind, 1.5, fc)
    be = mo["btri"] if dim == 3 else mo["bedges"]
    onbc = ffo.mark_bc(mo["nv"], be, mo["blab"], dim, labels)
    X = (mo["xyz"] if dim == 3 else mo["xy"]).reshape(-1, dim)
    gv = np.ascontiguousarray(X[:, 0] + 2 * X[:, 1])  # g at the vertices
    bo[onbc != 0] = 1e30 * gv[onbc != 0]
    b = ffhip.rhs_p1_form(m, kind, 1.5, fc, labels, 1e30, 0.0, gv)
    big = onbc != 0
    assert np.array_equal(b[big], bo[big])
    assert np.abs(b - bo)[~big].max() <= 1e-12 * np.abs(bo[~big]).max()


def test_rhs_form_device_arrays_and_constant_path():
    import torch
    mo, m = ffo.cube(6), ffhip.cube(6)
    fc = _qp(mo, 3)
    bo = ffo.rhs_p1_coef(mo, ffo.COEF_QP, 0.0, fc)
    onbc = ffo.mark_bc(mo["nv"], mo["btri"], mo["blab"], 3, (2, 5))
    ffo.rhs_bc(bo, onbc, 1e30, 0.75)
    out = torch.zeros(mo["nv"], dtype=torch.float64, device="cuda")
    ffhip.rhs_p1_form(m, ffhip.COEF_QP, 0.0, torch.tensor(fc.ravel(), device="cuda"), (2, 5), 1e30,
                      0.75, None, out)
    b = out.cpu().numpy()
    big = onbc != 0
    assert np.array_equal(b[big], bo[big])
    assert np.abs(b - bo)[~big].max() <= 1e-12 * np.abs(bo[~big]).max()
    # the constant-f entry point is the same kernel with kind CONST
    b1 = ffhip.rhs_p1(m, 2.0, (1, 2, 3, 4, 5, 6), 1e30, 0.0)
    b2 = ffhip.rhs_p1_form(m, ffhip.COEF_CONST, 2.0, None, (1, 2, 3, 4, 5, 6), 1e30, 0.0, None)
    assert np.array_equal(b1, b2)
