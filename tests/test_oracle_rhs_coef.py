"""Pins the oracle's linear form with a coefficient, int3d(Th)(f*v) (int2d in
2D): Element_rhs, problem.cpp:7940-7978, with f at the Tet_5 / T_5 points.

* constant f: equal bit for bit to the constant-f restatement, and its sum is
  f times the mesh measure;
* f = x at the quadrature points: the exact moments int_K x l_i =
  sum_k x_k |K| (1 + delta_ik) / 20 (tet), / 12 (triangle) — Tet_5 and T_5 are
  exact for degree 2, so a wrong point, weight or basis value fails;
* P0 f: each element adds f_e |K| / 4 (/ 3) to its vertices;
* the analytic coefficient through the element mapping equals the
  quadrature-point kind fed the same function.
"""
import numpy as np
import pytest

import ffo

MESHES = {"cube4": (3, lambda: ffo.cube(4)), "cube_2x3x4": (3, lambda: ffo.cube(2, 3, 4)),
          "square7x5": (2, lambda: ffo.square(7, 5))}


def _geom(m, dim):
    X = (m["xyz"] if dim == 3 else m["xy"]).reshape(-1, dim)
    el = (m["tets"] if dim == 3 else m["tris"]).reshape(-1, dim + 1)
    P = X[el]
    if dim == 3:
        mes = np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])) / 6
    else:
        e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
        mes = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    return X, el, P, mes


def _points(m, dim):
    X, el, P, mes = _geom(m, dim)
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
    return np.einsum("qi,eik->eqk", lam, P)


@pytest.mark.parametrize("name", list(MESHES))
def test_constant(name):
    dim, mk = MESHES[name]
    m = mk()
    b = ffo.rhs_p1_coef(m, ffo.COEF_CONST, 2.5)
    b0 = ffo.rhs_p1_3d(m, 2.5) if dim == 3 else ffo.rhs_p1_2d(m, 2.5)
    assert np.array_equal(b, b0)
    assert b.sum() == pytest.approx(2.5, rel=1e-13)  # unit cube / square


@pytest.mark.parametrize("name", list(MESHES))
def test_linear_f_exact_moments(name):
    dim, mk = MESHES[name]
    m = mk()
    X, el, P, mes = _geom(m, dim)
    fq = np.ascontiguousarray(_points(m, dim)[:, :, 0])
    b = ffo.rhs_p1_coef(m, ffo.COEF_QP, 0.0, fq)
    den = 20.0 if dim == 3 else 12.0
    ex = np.zeros(m["nv"])
    for e in range(el.shape[0]):
        xs = P[e, :, 0]
        for i in range(dim + 1):
            ex[el[e, i]] += mes[e] * (xs.sum() + xs[i]) / den
    assert np.abs(b - ex).max() <= 1e-14 * np.abs(ex).max()


@pytest.mark.parametrize("name", list(MESHES))
def test_p0_and_sin3(name):
    dim, mk = MESHES[name]
    m = mk()
    X, el, P, mes = _geom(m, dim)
    fe = np.random.default_rng(2).uniform(0.5, 2.0, el.shape[0])
    b = ffo.rhs_p1_coef(m, ffo.COEF_P0, 0.0, fe)
    ex = np.zeros(m["nv"])
    np.add.at(ex, el.ravel(), np.repeat(fe * mes / (dim + 1), dim + 1))
    assert np.abs(b - ex).max() <= 1e-14 * np.abs(ex).max()
    Q = _points(m, dim)
    tp = 2 * np.pi
    s = 0.5 * np.sin(tp * Q[..., 0]) * np.sin(tp * Q[..., 1])
    if dim == 3:
        s = s * np.sin(tp * Q[..., 2])
    b1 = ffo.rhs_p1_coef(m, ffo.COEF_SIN3)
    b2 = ffo.rhs_p1_coef(m, ffo.COEF_QP, 0.0, np.ascontiguousarray(1 + s))
    assert np.abs(b1 - b2).max() <= 1e-14 * np.abs(b2).max()
