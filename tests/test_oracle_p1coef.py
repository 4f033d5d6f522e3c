"""The oracle's P1 evaluation at the quadrature points (ffo_p1_at_qp, the
restatement of FreeFEM's TypeOfFE_P1Lagrange3d::operator(), P012_3d.cpp:98-106,
and TypeOfFE_P1Lagrange::operator(), FESpace.cpp:1078-1085), pinned by what
the mathematics fixes: a P1 function that is affine on the whole mesh is its
own interpolant, so its value at each Tet_5 / T_5 point equals the affine
function at the mapped point x_q = sum_i l_i(q) X_i (independently computed
here from the quadrature table); vertex values (a point at a vertex) and the
partition of unity (u = 1) follow.  A wrong barycentric (a swapped l_i, a
dropped term) fails the affine check at all but symmetric points.
"""
import numpy as np
import pytest

import ffo


def _points(dim):
    if dim == 3:
        p, w = np.zeros(42), np.zeros(14)
        ffo.lib().ffo_qf_tet5(p, w)
        return p.reshape(14, 3)
    p, w = np.zeros(14), np.zeros(7)
    ffo.lib().ffo_qf_tri5(p, w)
    return p.reshape(7, 2)


@pytest.mark.parametrize("dim", [3, 2])
def test_affine_function_is_reproduced(dim):
    m = ffo.cube(3, 4, 2) if dim == 3 else ffo.square(5, 3)
    X = (m["xyz"] if dim == 3 else m["xy"]).reshape(-1, dim)
    el = (m["tets"] if dim == 3 else m["tris"]).reshape(-1, dim + 1)
    coef = np.array([0.7, -1.3, 2.1, 0.4][:dim + 1])
    u = coef[0] + X @ coef[1:]
    got = ffo.p1_at_qp(m, u).reshape(len(el), -1)
    P = _points(dim)
    lam = np.concatenate([1 - P.sum(1, keepdims=True), P], axis=1)  # nq x (dim+1)
    Q = np.einsum("qi,eik->eqk", lam, X[el])
    want = coef[0] + Q @ coef[1:]
    assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want))
    ones = ffo.p1_at_qp(m, np.ones(len(X)))
    assert np.max(np.abs(ones - 1)) <= 4e-16


def test_random_function_is_the_barycentric_sum():
    m = ffo.cube(2)
    u = np.random.default_rng(4).standard_normal(m["nv"])
    el = m["tets"].reshape(-1, 4)
    P = _points(3)
    lam = np.concatenate([1 - P.sum(1, keepdims=True), P], axis=1)
    want = np.einsum("qi,ei->eq", lam, u[el])
    got = ffo.p1_at_qp(m, u).reshape(len(el), 14)
    assert np.max(np.abs(got - want)) <= 1e-15 * np.max(np.abs(want)) * 8
